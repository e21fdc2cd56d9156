// Batches of structures of uneven size (node, mode, component and frequency
// counts): every structure is a resident plan; one step runs all their tables
// (spread over worker streams), then ONE persistent G-contraction launch whose
// workgroups take work items (structure, tile pair, mode) from a device queue,
// while the worker streams compute the modal PSD moments.  The step is
// captured once in a HIP graph (forked and joined streams become parallel
// branches) and replayed.  The item list is ordered by decreasing cost per
// item (longest-processing-time first) and split into one queue per XCD; a
// small autotuner picks the persistent grid size and the items per ticket.

__device__ __forceinline__ long uniform_long(long v)
{
   const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffffL));
   const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
   return ((long)hi << 32) | (long)(unsigned)lo;
}

// structure of item `it`: offs[s] <= it < offs[s+1] (offs has ns+1 entries)
__device__ __forceinline__ int batch_struct(const long *__restrict__ offs, int ns, long it)
{
   int lo = 0, hi = ns;
   while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (offs[mid] <= it) lo = mid;
      else hi = mid;
   }
   return lo;
}

__global__ void k_zero_u64(unsigned long long *c, int n)
{
   if ((int)threadIdx.x < n) c[threadIdx.x] = 0ull;
}

// Work queues: the ordered item list is cut into 8 contiguous slices, one per
// XCD (a persistent workgroup stays on XCD blockIdx.x % 8), so the nm items of
// one tile pair -- which read the same U tiles -- share that XCD's L2.  A
// workgroup drains its own slice, then takes items from the other slices.
#define BATCH_QUEUES 8

template <class Item>
__device__ __forceinline__ void batch_drain(const long *__restrict__ offs, int ns,
                                            const long *__restrict__ qbeg,
                                            unsigned long long *counters, int chunk, Item item)
{
   __shared__ long s_i0;
   const int own = blockIdx.x % BATCH_QUEUES;
   for (int k = 0; k < BATCH_QUEUES; ++k) {
      const int q = (own + k) % BATCH_QUEUES;
      const long b0 = qbeg[q], len = qbeg[q + 1] - qbeg[q];
      for (;;) {
         __syncthreads();
         if (threadIdx.x == 0) s_i0 = (long)atomicAdd(counters + q, (unsigned long long)chunk);
         __syncthreads();
         const long i0 = uniform_long(s_i0);
         if (i0 >= len) break;
         const long i1 = i0 + chunk < len ? i0 + chunk : len;
         for (long it = b0 + i0; it < b0 + i1; ++it) {
            const int s = batch_struct(offs, ns, it);
            item(s, it - offs[s]);
         }
      }
   }
}

__global__ __launch_bounds__(256, 4) void k_gsym_f64_batch(const GsymArgs *__restrict__ args,
                                                        const long *__restrict__ offs, int ns,
                                                        const long *__restrict__ qbeg,
                                                        unsigned long long *counters, int chunk)
{
   batch_drain(offs, ns, qbeg, counters, chunk, [&](int s, long li) {
      const GsymArgs a = args[s];
      gsym_f64_item(a, (int)(li % a.nm), li / a.nm);
   });
}

__global__ __launch_bounds__(256, 2) void k_gsym_bf16x3_batch(
   const GsymArgs *__restrict__ args, const GsymArgs16 *__restrict__ args16,
   const long *__restrict__ offs, int ns, const long *__restrict__ qbeg,
   unsigned long long *counters, int chunk)
{
   batch_drain(offs, ns, qbeg, counters, chunk, [&](int s, long li) {
      const GsymArgs a = args[s];
      const GsymArgs16 b = args16[s];
      gsym_bf16x3_item(a, b, (int)(li % a.nm), li / a.nm);
   });
}

__global__ __launch_bounds__(256, 2) void k_gsym_f32_batch(
   const GsymArgs *__restrict__ args, const GsymArgs16 *__restrict__ args16,
   const long *__restrict__ offs, int ns, const long *__restrict__ qbeg,
   unsigned long long *counters, int chunk)
{
   batch_drain(offs, ns, qbeg, counters, chunk, [&](int s, long li) {
      const GsymArgs a = args[s];
      const GsymArgs16 b = args16[s];
      gsym_f32_item(a, b, (int)(li % a.nm), li / a.nm);
   });
}

// results of every structure into one contiguous buffer: array j = srcs[j]
// (lens[j] doubles) to dst + offs[j]; one block per array
__global__ void k_pack(const double *const *__restrict__ srcs, const long *__restrict__ lens,
                       const long *__restrict__ offs, double *__restrict__ dst)
{
   const int j = blockIdx.x;
   const double *src = srcs[j];
   double *d = dst + offs[j];
   for (long t = threadIdx.x; t < lens[j]; t += blockDim.x) d[t] = src[t];
}

struct bsahip_batch {
   std::vector<std::unique_ptr<bsahip_plan>> plans;
   std::vector<int> order;            // queue order of the structures
   int prec = 0;
   int32_t flags = 0;
   Bufs B;
   GsymArgs *args = nullptr;
   GsymArgs16 *args16 = nullptr;
   long *offs = nullptr;
   unsigned long long *counter = nullptr;   // one per queue
   long *qbeg = nullptr;                    // queue slices of the item list
   std::vector<long> qb;                    // host copy
   long total = 0;
   int ncu = 0, occ = 1;
   int grid_mult = 2, chunk = 1;
   hipGraph_t graph = nullptr;
   hipGraphExec_t gexec = nullptr;
   double ms_total = 0.0, ms_gsym = 0.0;
   // worker streams (default priority) for the per-structure table and PSD
   // moment kernels, forked from and joined to the library stream
   std::vector<hipStream_t> ws;
   hipEvent_t ev_fork = nullptr;
   std::vector<hipEvent_t> ev_prep, ev_m2;
   // packed results: per structure m3 (m3f, m3r) then m2 (f, r, o)
   const double **psrc = nullptr;
   long *plen = nullptr, *poff = nullptr;
   double *pack = nullptr, *hpack = nullptr;   // device buffer, pinned host copy
   std::vector<long> m3off, m2off;
   long npack = 0;
   ~bsahip_batch()
   {
      if (hpack) (void)hipHostFree(hpack);
      if (gexec) (void)hipGraphExecDestroy(gexec);
      if (graph) (void)hipGraphDestroy(graph);
      for (auto e : ev_prep) (void)hipEventDestroy(e);
      for (auto e : ev_m2) (void)hipEventDestroy(e);
      if (ev_fork) (void)hipEventDestroy(ev_fork);
      for (auto st : ws) (void)hipStreamDestroy(st);
   }
};

namespace {

// queue slices for the first `total` items: 8 equal contiguous ranges
int batch_slices(bsahip_batch *b, long total)
{
   b->qb.assign(BATCH_QUEUES + 1, 0);
   for (int q = 0; q <= BATCH_QUEUES; ++q) b->qb[q] = total * q / BATCH_QUEUES;
   HIPC(hipMemcpyAsync(b->qbeg, b->qb.data(), (BATCH_QUEUES + 1) * sizeof(long),
                       hipMemcpyHostToDevice, g.s));
   return 0;
}

int batch_launch(bsahip_batch *b)
{
   const unsigned grid = (unsigned)std::max(BATCH_QUEUES, b->grid_mult * b->ncu * b->occ);
   const int ns = (int)b->order.size();
   k_zero_u64<<<1, 64, 0, g.s>>>(b->counter, BATCH_QUEUES);
   HIPC(hipGetLastError());
   if (b->prec == BSAHIP_PREC_BF16X3)
      k_gsym_bf16x3_batch<<<grid, 256, 0, g.s>>>(b->args, b->args16, b->offs, ns, b->qbeg,
                                                 b->counter, b->chunk);
   else if (b->prec == BSAHIP_PREC_FP32)
      k_gsym_f32_batch<<<grid, 256, 0, g.s>>>(b->args, b->args16, b->offs, ns, b->qbeg,
                                              b->counter, b->chunk);
   else
      k_gsym_f64_batch<<<grid, 256, 0, g.s>>>(b->args, b->offs, ns, b->qbeg, b->counter, b->chunk);
   HIPC(hipGetLastError());
   return 0;
}

// the library stream is g.s; run f with g.s switched to another stream
template <class F>
int on_stream(hipStream_t st, F f)
{
   const hipStream_t keep = g.s;
   g.s = st;
   const int rc = f();
   g.s = keep;
   return rc;
}

// one step, stream-ordered, no host synchronisation (capturable).
// mode 1: everything on the library stream, one G launch per structure.
// mode 0: tables on the worker streams, then the persistent G launch on the
// library stream while the workers compute the PSD moments.
int batch_step(bsahip_batch *b, int mode)
{
   const bool bisp = !(b->flags & BSAHIP_NO_BISP), psd = !(b->flags & BSAHIP_NO_PSD);
   if (mode == 1) {
      for (auto &P : b->plans) CHK(plan_prepare(P.get()));
      if (bisp)
         for (int s : b->order) CHK(plan_gsym(b->plans[s].get()));
      if (psd)
         for (auto &P : b->plans) CHK(plan_m2(P.get()));
      k_pack<<<(unsigned)(2 * b->plans.size()), 256, 0, g.s>>>(b->psrc, b->plen, b->poff, b->pack);
      HIPC(hipGetLastError());
      return 0;
   }
   const int nw = (int)b->ws.size();
   const hipStream_t main = g.s;
   HIPC(hipEventRecord(b->ev_fork, main));
   for (int w = 0; w < nw; ++w) HIPC(hipStreamWaitEvent(b->ws[w], b->ev_fork, 0));
   for (size_t i = 0; i < b->plans.size(); ++i)
      CHK(on_stream(b->ws[i % nw], [&] { return plan_prepare(b->plans[i].get()); }));
   for (int w = 0; w < nw; ++w) {
      HIPC(hipEventRecord(b->ev_prep[w], b->ws[w]));
      HIPC(hipStreamWaitEvent(main, b->ev_prep[w], 0));
   }
   if (bisp) CHK(batch_launch(b));
   if (psd)
      for (size_t i = 0; i < b->plans.size(); ++i)
         CHK(on_stream(b->ws[i % nw], [&] { return plan_m2(b->plans[i].get()); }));
   for (int w = 0; w < nw; ++w) {
      HIPC(hipEventRecord(b->ev_m2[w], b->ws[w]));
      HIPC(hipStreamWaitEvent(main, b->ev_m2[w], 0));
   }
   k_pack<<<(unsigned)(2 * b->plans.size()), 256, 0, main>>>(b->psrc, b->plen, b->poff, b->pack);
   HIPC(hipGetLastError());
   return 0;
}

void batch_drop_graph(bsahip_batch *b)
{
   if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
   if (b->graph) (void)hipGraphDestroy(b->graph);
   b->gexec = nullptr;
   b->graph = nullptr;
}

int batch_capture(bsahip_batch *b)
{
   batch_drop_graph(b);
   HIPC(hipStreamBeginCapture(g.s, hipStreamCaptureModeRelaxed));
   const int rc = batch_step(b, 0);
   hipGraph_t gr = nullptr;
   const hipError_t e = hipStreamEndCapture(g.s, &gr);
   if (rc) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
   }
   if (e != hipSuccess) return fail(BSAHIP_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e));
   b->graph = gr;
   HIPC(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
   return 0;
}

}  // namespace

extern "C" {

int bsahip_batch_create(const bsahip_problem *const *probs, const int32_t *nfh, const double *df,
                        int32_t n, int32_t flags, bsahip_batch **out)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!probs || !nfh || !df || !out || n < 1) return fail(BSAHIP_EINVAL, "bad batch arguments");
   *out = nullptr;
   CHK(ensure_init());
   CHK(check_flags(flags, true, true));
   if (flags & BSAHIP_SUM_EXACT) return fail(BSAHIP_EUNSUPPORTED, "batches use the Classic truncated sum");
   if (flags & BSAHIP_DETERMINISTIC)
      return fail(BSAHIP_EUNSUPPORTED, "batches sum with atomics (no fixed-order mode)");
   std::unique_ptr<bsahip_batch> B(new bsahip_batch());
   B->prec = flags & BSAHIP_PREC_MASK;
   B->flags = flags;
   for (int i = 0; i < n; ++i) {
      std::unique_ptr<bsahip_plan> P(new bsahip_plan());
      const int rc = plan_setup(P.get(), probs[i], nfh[i], df[i], flags, 0);
      if (rc) {
         g_err = "structure " + std::to_string(i) + ": " + g_err;
         return rc;
      }
      B->plans.push_back(std::move(P));
   }
   // queue order: decreasing contraction length K (cost of one item), then modes
   B->order.resize(n);
   for (int i = 0; i < n; ++i) B->order[i] = i;
   std::stable_sort(B->order.begin(), B->order.end(), [&](int x, int y) {
      const Prob &a = B->plans[x]->D, &c = B->plans[y]->D;
      return a.K != c.K ? a.K > c.K : a.nm > c.nm;
   });
   std::vector<GsymArgs> a(n);
   std::vector<GsymArgs16> a16(n);
   std::vector<long> offs(n + 1, 0);
   for (int q = 0; q < n; ++q) {
      const bsahip_plan *P = B->plans[B->order[q]].get();
      plan_args(P, a[q], a16[q]);
      offs[q + 1] = offs[q] + plan_items(P);
   }
   B->total = offs[n];
   CHK(B->B.get("args", (size_t)n, &B->args));
   CHK(B->B.get("args16", (size_t)n, &B->args16));
   CHK(B->B.get("offs", (size_t)n + 1, &B->offs));
   CHK(B->B.get("counter", BATCH_QUEUES, &B->counter));
   CHK(B->B.get("qbeg", BATCH_QUEUES + 1, &B->qbeg));
   HIPC(hipMemcpyAsync(B->args, a.data(), n * sizeof(GsymArgs), hipMemcpyHostToDevice, g.s));
   HIPC(hipMemcpyAsync(B->args16, a16.data(), n * sizeof(GsymArgs16), hipMemcpyHostToDevice, g.s));
   HIPC(hipMemcpyAsync(B->offs, offs.data(), (n + 1) * sizeof(long), hipMemcpyHostToDevice, g.s));
   CHK(batch_slices(B.get(), B->total));
   int dev = 0;
   HIPC(hipGetDevice(&dev));
   HIPC(hipDeviceGetAttribute(&B->ncu, hipDeviceAttributeMultiprocessorCount, dev));
   int occ = 1;
   if (B->prec == BSAHIP_PREC_BF16X3)
      HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_gsym_bf16x3_batch, 256, 0));
   else if (B->prec == BSAHIP_PREC_FP32)
      HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_gsym_f32_batch, 256, 0));
   else
      HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_gsym_f64_batch, 256, 0));
   B->occ = std::max(1, occ);
   const int nw = std::min(8, n);
   B->ws.resize(nw, nullptr);
   B->ev_prep.resize(nw, nullptr);
   B->ev_m2.resize(nw, nullptr);
   for (int w = 0; w < nw; ++w) {
      HIPC(hipStreamCreateWithFlags(&B->ws[w], hipStreamNonBlocking));
      HIPC(hipEventCreateWithFlags(&B->ev_prep[w], hipEventDisableTiming));
      HIPC(hipEventCreateWithFlags(&B->ev_m2[w], hipEventDisableTiming));
   }
   HIPC(hipEventCreateWithFlags(&B->ev_fork, hipEventDisableTiming));
   {
      std::vector<const double *> src;
      std::vector<long> len, off;
      long o = 0;
      for (auto &P : B->plans) {
         const long nm = P->D.nm;
         B->m3off.push_back(o);
         src.push_back(P->m3);
         len.push_back(2 * nm * nm * nm);
         off.push_back(o);
         o += 2 * nm * nm * nm;
         B->m2off.push_back(o);
         src.push_back(P->m2);
         len.push_back(3 * nm * nm);
         off.push_back(o);
         o += 3 * nm * nm;
      }
      B->npack = o;
      CHK(B->B.get("psrc", src.size(), &B->psrc));
      CHK(B->B.get("plen", len.size(), &B->plen));
      CHK(B->B.get("poff", off.size(), &B->poff));
      CHK(B->B.get("pack", (size_t)o, &B->pack));
      HIPC(hipMemcpyAsync(B->psrc, src.data(), src.size() * sizeof(double *), hipMemcpyHostToDevice, g.s));
      HIPC(hipMemcpyAsync(B->plen, len.data(), len.size() * sizeof(long), hipMemcpyHostToDevice, g.s));
      HIPC(hipMemcpyAsync(B->poff, off.data(), off.size() * sizeof(long), hipMemcpyHostToDevice, g.s));
      HIPC(hipHostMalloc((void **)&B->hpack, (size_t)o * sizeof(double), hipHostMallocDefault));
   }
   // warm-up run: every buffer reaches its size before a graph is captured
   CHK(batch_step(B.get(), 1));
   HIPC(hipStreamSynchronize(g.s));
   *out = B.release();
   return 0;
}

int bsahip_batch_autotune(bsahip_batch *b, int32_t *grid_mult, int32_t *chunk)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!b) return fail(BSAHIP_EINVAL, "NULL batch");
   if (b->flags & BSAHIP_NO_BISP) return 0;
   // the tables of every structure are current after any step; time the
   // persistent contraction alone on the first quarter of the queue (the
   // structures with the longest items)
   const long part = std::max(1L, b->total / 4);
   static const int mults[] = {1, 2, 4};
   static const int chunks[] = {1, 4, 16};
   double best = 1e300;
   int bm = b->grid_mult, bc = b->chunk;
   for (int m : mults)
      for (int c : chunks) {
         b->grid_mult = m;
         b->chunk = c;
         CHK(batch_slices(b, part));
         HIPC(hipEventRecord(g.e0, g.s));
         CHK(batch_launch(b));
         HIPC(hipEventRecord(g.e1, g.s));
         HIPC(hipStreamSynchronize(g.s));
         float ms = 0.f;
         HIPC(hipEventElapsedTime(&ms, g.e0, g.e1));
         if (ms < best) {
            best = ms;
            bm = m;
            bc = c;
         }
      }
   b->grid_mult = bm;
   b->chunk = bc;
   CHK(batch_slices(b, b->total));
   HIPC(hipStreamSynchronize(g.s));
   batch_drop_graph(b);
   if (grid_mult) *grid_mult = bm;
   if (chunk) *chunk = bc;
   return 0;
}

int bsahip_batch_run(bsahip_batch *b, int32_t mode, double *const *m3mf, double *const *m3mr,
                     double *const *m2mr)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!b) return fail(BSAHIP_EINVAL, "NULL batch");
   if (mode < 0 || mode > 2)
      return fail(BSAHIP_EINVAL, "mode must be 0 (graph), 1 (per structure) or 2 (streams)");
   if (mode == 0 && !b->gexec) CHK(batch_capture(b));
   HIPC(hipEventRecord(g.e0, g.s));
   if (mode == 0) HIPC(hipGraphLaunch(b->gexec, g.s));
   else CHK(batch_step(b, mode == 1 ? 1 : 0));
   HIPC(hipEventRecord(g.e1, g.s));
   HIPC(hipMemcpyAsync(b->hpack, b->pack, (size_t)b->npack * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   for (size_t i = 0; i < b->plans.size(); ++i) {
      const size_t nm = b->plans[i]->D.nm, nm3 = nm * nm * nm, nm2 = nm * nm;
      const double *m3 = b->hpack + b->m3off[i], *m2 = b->hpack + b->m2off[i];
      if (m3mf && m3mf[i]) std::memcpy(m3mf[i], m3, nm3 * sizeof(double));
      if (m3mr && m3mr[i]) std::memcpy(m3mr[i], m3 + nm3, nm3 * sizeof(double));
      if (m2mr && m2mr[i]) std::memcpy(m2mr[i], m2 + nm2, nm2 * sizeof(double));
   }
   float ms = 0.f;
   HIPC(hipEventElapsedTime(&ms, g.e0, g.e1));
   b->ms_total = ms;
   return 0;
}

int bsahip_batch_info(const bsahip_batch *b, double *ms_step, int64_t *items, int32_t *grid_mult,
                      int32_t *chunk)
{
   if (!b) return fail(BSAHIP_EINVAL, "NULL batch");
   if (ms_step) *ms_step = b->ms_total;
   if (items) *items = b->total;
   if (grid_mult) *grid_mult = b->grid_mult;
   if (chunk) *chunk = b->chunk;
   return 0;
}

int bsahip_batch_destroy(bsahip_batch *b)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   delete b;
   return 0;
}

}  // extern "C"
