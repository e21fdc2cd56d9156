// Multilayer perceptron of the learned Mesher proposer: trained and evaluated
// on the device by the kernels below (FP32 weights and activations).
//
//   z_l = W_l a_{l-1} + b_l,  a_l = GELU(z_l) for hidden layers, logit = z_L
//   GELU(x) = x Phi(x) = x (1 + erf(x/sqrt 2)) / 2,  GELU'(x) = Phi(x) + x phi(x)
//   loss = sum_i w_i [ y_i softplus(-z_i) + (1 - y_i) softplus(z_i) ] / sum_i w_i
//   (binary cross-entropy on logits), Adam (Kingma & Ba 2015) with bias correction.
//
// Inputs are normalised on the device with the stored per-feature mean and
// scale before the first layer.  Activations are row-major [batch][features];
// W_l is [out][in].

// C[i*ldc + j] = sum_k A(i,k) B(k,j) (+ bias[j]), A(i,k) = A[i*sai + k*sak],
// B(k,j) = B[k*sbk + j*sbj]; 64x64 tile per 256 threads (4x4 outputs each),
// K staged by 16 in LDS.  act: 0 none, 1 store z in Z and GELU(z) in C.
__global__ __launch_bounds__(256) void k_mlp_gemm(
   int M, int N, int K, const float *__restrict__ A, long sai, long sak,
   const float *__restrict__ B, long sbk, long sbj, const float *__restrict__ bias,
   float *__restrict__ C, long ldc, float *__restrict__ Z, int act, int accumulate)
{
   __shared__ float sa[16][64 + 4];
   __shared__ float sb[16][64 + 4];
   const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
   const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
   float acc[4][4];
#pragma unroll
   for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;
   for (int k0 = 0; k0 < K; k0 += 16) {
      for (int t = threadIdx.x; t < 16 * 64; t += 256) {
         const int kk = t >> 6, r = t & 63;
         const int i = i0 + r, k = k0 + kk, j = j0 + r;
         sa[kk][r] = (i < M && k < K) ? A[(size_t)i * sai + (size_t)k * sak] : 0.f;
         sb[kk][r] = (j < N && k < K) ? B[(size_t)k * sbk + (size_t)j * sbj] : 0.f;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) {
         float av[4], bv[4];
#pragma unroll
         for (int a = 0; a < 4; ++a) av[a] = sa[kk][ty + 16 * a];
#pragma unroll
         for (int b = 0; b < 4; ++b) bv[b] = sb[kk][tx + 16 * b];
#pragma unroll
         for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fmaf(av[a], bv[b], acc[a][b]);
      }
      __syncthreads();
   }
#pragma unroll
   for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
         const int i = i0 + ty + 16 * a, j = j0 + tx + 16 * b;
         if (i >= M || j >= N) continue;
         float z = acc[a][b] + (bias ? bias[j] : 0.f);
         float *c = C + (size_t)i * ldc + j;
         if (act == 1) {
            Z[(size_t)i * ldc + j] = z;
            z = 0.5f * z * (1.f + erff(z * 0.70710678118654752f));
         }
         *c = accumulate ? *c + z : z;
      }
}

// x_n = (x - mean) * inv_scale, row-major [B][F]
__global__ void k_mlp_normalise(const float *__restrict__ x, int Bn, int F,
                                const float *__restrict__ mean, const float *__restrict__ inv,
                                float *__restrict__ out)
{
   const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (gid >= (long)Bn * F) return;
   const int f = (int)(gid % F);
   out[gid] = (x[gid] - mean[f]) * inv[f];
}

// loss terms and dL/dz for the output logits (N = 1 column)
__global__ void k_mlp_bce(const float *__restrict__ z, const float *__restrict__ y,
                          const float *__restrict__ w, int Bn, float inv_wsum,
                          float *__restrict__ dz, float *__restrict__ loss_terms)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= Bn) return;
   const float zi = z[i], yi = y[i], wi = w[i];
   // softplus(t) = max(t,0) + log1p(exp(-|t|))
   const float sp_pos = fmaxf(zi, 0.f) + log1pf(expf(-fabsf(zi)));     // softplus(z)
   const float sp_neg = sp_pos - zi;                                   // softplus(-z)
   loss_terms[i] = wi * (yi * sp_neg + (1.f - yi) * sp_pos) * inv_wsum;
   const float sg = 1.f / (1.f + expf(-zi));
   dz[i] = wi * (sg - yi) * inv_wsum;
}

// dZ = dA * GELU'(Z)
__global__ void k_mlp_gelu_back(const float *__restrict__ Z, float *__restrict__ dA, long n)
{
   const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   const float z = Z[i];
   const float cdf = 0.5f * (1.f + erff(z * 0.70710678118654752f));
   const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
   dA[i] *= cdf + z * pdf;
}

// db[j] = sum_i dZ[i][j]  (one block per column chunk)
__global__ void k_mlp_colsum(const float *__restrict__ dZ, int Bn, int N, float *__restrict__ db)
{
   const int j = blockIdx.x;
   float s = 0.f;
   for (int i = threadIdx.x; i < Bn; i += blockDim.x) s += dZ[(size_t)i * N + j];
   __shared__ float red[256];
   red[threadIdx.x] = s;
   __syncthreads();
   for (int o = blockDim.x / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
   }
   if (threadIdx.x == 0) db[j] = red[0];
}

__global__ void k_mlp_sum(const float *__restrict__ x, int n, float *__restrict__ out)
{
   __shared__ float red[256];
   float s = 0.f;
   for (int i = threadIdx.x; i < n; i += 256) s += x[i];
   red[threadIdx.x] = s;
   __syncthreads();
   for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
   }
   if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void k_mlp_adam(float *__restrict__ p, const float *__restrict__ g,
                           float *__restrict__ m, float *__restrict__ v, long n, float lr,
                           float b1, float b2, float eps, float c1, float c2, float wd)
{
   const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   const float gi = g[i] + wd * p[i];
   const float mi = b1 * m[i] + (1.f - b1) * gi;
   const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
   m[i] = mi;
   v[i] = vi;
   p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// gather rows idx[r] of X [n][F] into out [B][F] (minibatch), and y, w
__global__ void k_mlp_gather(const float *__restrict__ X, const float *__restrict__ y,
                             const float *__restrict__ w, const int *__restrict__ idx, int Bn,
                             int F, float *__restrict__ Xo, float *__restrict__ yo,
                             float *__restrict__ wo)
{
   const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (gid >= (long)Bn * F) return;
   const int r = (int)(gid / F), f = (int)(gid % F);
   const int s = idx[r];
   Xo[gid] = X[(size_t)s * F + f];
   if (f == 0) {
      yo[r] = y[s];
      wo[r] = w[s];
   }
}

// ===========================================================================
// host side
// ===========================================================================
struct bsahip_mlp {
   Bufs B;
   int nin = 0, width = 0, depth = 0;      // depth = number of linear layers
   std::vector<int> din, dout;             // per layer
   std::vector<size_t> woff, boff;         // offsets into the flat parameter vector
   size_t nparam = 0;
   float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;   // params, grads, Adam
   float *mean = nullptr, *inv = nullptr;
   long step = 0;
};

namespace {

int mlp_gemm(int M, int N, int K, const float *A, long sai, long sak, const float *B, long sbk,
             long sbj, const float *bias, float *C, long ldc, float *Z, int act, int accumulate)
{
   dim3 gs((N + 63) / 64, (M + 63) / 64);
   k_mlp_gemm<<<gs, 256, 0, g.s>>>(M, N, K, A, sai, sak, B, sbk, sbj, bias, C, ldc, Z, act,
                                   accumulate);
   HIPC(hipGetLastError());
   return 0;
}

// forward over Bn normalised rows already in device X0 [Bn][nin]; keeps the
// pre-activations Z_l and activations A_l (l = 0..depth-1) in named buffers.
int mlp_forward(bsahip_mlp *N, const float *X0, int Bn, float **logits)
{
   const float *a = X0;
   for (int l = 0; l < N->depth; ++l) {
      const int di = N->din[l], dn = N->dout[l];
      float *Z, *Aout;
      const std::string zs = "Z" + std::to_string(l), as = "A" + std::to_string(l);
      CHK(N->B.get(zs.c_str(), (size_t)Bn * dn, &Z));
      CHK(N->B.get(as.c_str(), (size_t)Bn * dn, &Aout));
      const bool hidden = l + 1 < N->depth;
      // A (Bn x di) row-major; W [dn][di] used as B(k=i, j=o) = W[o*di + i]
      CHK(mlp_gemm(Bn, dn, di, a, di, 1, N->P + N->woff[l], 1, di, N->P + N->boff[l],
                   hidden ? Aout : Z, dn, Z, hidden ? 1 : 0, 0));
      a = hidden ? Aout : Z;
   }
   *logits = const_cast<float *>(a);
   return 0;
}

// loss and gradients (into N->G) for normalised rows X0 with labels y, weights w
int mlp_backward(bsahip_mlp *N, const float *X0, const float *y, const float *w, int Bn,
                 float wsum, float *loss_dev)
{
   float *z;
   CHK(mlp_forward(N, X0, Bn, &z));
   float *dz, *lt;
   CHK(N->B.get("dz_out", (size_t)Bn, &dz));
   CHK(N->B.get("lterms", (size_t)Bn, &lt));
   k_mlp_bce<<<(Bn + 255) / 256, 256, 0, g.s>>>(z, y, w, Bn, 1.f / wsum, dz, lt);
   HIPC(hipGetLastError());
   k_mlp_sum<<<1, 256, 0, g.s>>>(lt, Bn, loss_dev);
   HIPC(hipGetLastError());
   float *dcur = dz;               // dL/dz of layer l, [Bn][dout]
   for (int l = N->depth - 1; l >= 0; --l) {
      const int di = N->din[l], dn = N->dout[l];
      const float *ain = (l == 0) ? X0 : nullptr;
      if (l > 0) {
         const std::string as = "A" + std::to_string(l - 1);
         float *ap;
         CHK(N->B.get(as.c_str(), (size_t)Bn * di, &ap));
         ain = ap;
      }
      // dW[o][i] = sum_b dz[b][o] a[b][i]: M = dn, N = di, K = Bn
      CHK(mlp_gemm(dn, di, Bn, dcur, 1, dn, ain, di, 1, nullptr, N->G + N->woff[l], di, nullptr,
                   0, 0));
      k_mlp_colsum<<<dn, 256, 0, g.s>>>(dcur, Bn, dn, N->G + N->boff[l]);
      HIPC(hipGetLastError());
      if (l > 0) {
         // dA_{l-1}[b][i] = sum_o dz[b][o] W[o][i]; then times GELU'(Z_{l-1})
         float *da, *Zp;
         const std::string ds = "dA" + std::to_string(l - 1), zs = "Z" + std::to_string(l - 1);
         CHK(N->B.get(ds.c_str(), (size_t)Bn * di, &da));
         CHK(N->B.get(zs.c_str(), (size_t)Bn * di, &Zp));
         CHK(mlp_gemm(Bn, di, dn, dcur, dn, 1, N->P + N->woff[l], di, 1, nullptr, da, di, nullptr,
                      0, 0));
         const long n = (long)Bn * di;
         k_mlp_gelu_back<<<(unsigned)((n + 255) / 256), 256, 0, g.s>>>(Zp, da, n);
         HIPC(hipGetLastError());
         dcur = da;
      }
   }
   return 0;
}

int mlp_normalise(bsahip_mlp *N, const float *xd, int Bn, float **out, const char *name)
{
   CHK(N->B.get(name, (size_t)Bn * N->nin, out));
   const long n = (long)Bn * N->nin;
   k_mlp_normalise<<<(unsigned)((n + 255) / 256), 256, 0, g.s>>>(xd, Bn, N->nin, N->mean, N->inv,
                                                                  *out);
   HIPC(hipGetLastError());
   return 0;
}

}  // namespace

extern "C" {

int bsahip_mlp_create(int32_t nin, int32_t width, int32_t depth, uint64_t seed, bsahip_mlp **out)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!out || nin < 1 || width < 1 || depth < 1 || depth > 16)
      return fail(BSAHIP_EINVAL, "need nin, width >= 1 and 1 <= depth <= 16");
   *out = nullptr;
   CHK(ensure_init());
   std::unique_ptr<bsahip_mlp> N(new bsahip_mlp());
   N->nin = nin;
   N->width = width;
   N->depth = depth;
   size_t off = 0;
   for (int l = 0; l < depth; ++l) {
      N->din.push_back(l == 0 ? nin : width);
      N->dout.push_back(l + 1 == depth ? 1 : width);
      N->woff.push_back(off);
      off += (size_t)N->din[l] * N->dout[l];
      N->boff.push_back(off);
      off += N->dout[l];
   }
   N->nparam = off;
   // He-style initialisation, weights ~ U(-s, s) with s = sqrt(6 / fan_in), biases 0
   std::vector<float> p(off, 0.f);
   uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
   auto rnd = [&]() {
      st ^= st << 13;
      st ^= st >> 7;
      st ^= st << 17;
      return (double)(st >> 11) * (1.0 / 9007199254740992.0);
   };
   for (int l = 0; l < depth; ++l) {
      const double s = std::sqrt(6.0 / N->din[l]) * (l + 1 == depth ? 0.1 : 1.0);
      for (size_t q = 0; q < (size_t)N->din[l] * N->dout[l]; ++q)
         p[N->woff[l] + q] = (float)((2.0 * rnd() - 1.0) * s);
   }
   CHK(upload(N->B, "P", p, &N->P));
   CHK(N->B.get("G", off, &N->G));
   std::vector<float> z(off, 0.f);
   CHK(upload(N->B, "Mo", z, &N->Mo));
   CHK(upload(N->B, "Vo", z, &N->Vo));
   std::vector<float> mean(nin, 0.f), inv(nin, 1.f);
   CHK(upload(N->B, "mean", mean, &N->mean));
   CHK(upload(N->B, "inv", inv, &N->inv));
   HIPC(hipStreamSynchronize(g.s));
   *out = N.release();
   return 0;
}

int bsahip_mlp_nparams(const bsahip_mlp *N, int64_t *n)
{
   if (!N || !n) return fail(BSAHIP_EINVAL, "NULL argument");
   *n = (int64_t)N->nparam;
   return 0;
}

int bsahip_mlp_set_params(bsahip_mlp *N, const float *params, const float *mean,
                          const float *inv_scale)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!N) return fail(BSAHIP_EINVAL, "NULL MLP");
   if (params) HIPC(hipMemcpyAsync(N->P, params, N->nparam * 4, hipMemcpyHostToDevice, g.s));
   if (mean) HIPC(hipMemcpyAsync(N->mean, mean, (size_t)N->nin * 4, hipMemcpyHostToDevice, g.s));
   if (inv_scale) HIPC(hipMemcpyAsync(N->inv, inv_scale, (size_t)N->nin * 4, hipMemcpyHostToDevice, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

int bsahip_mlp_get_params(bsahip_mlp *N, float *params, float *grads)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!N) return fail(BSAHIP_EINVAL, "NULL MLP");
   if (params) HIPC(hipMemcpyAsync(params, N->P, N->nparam * 4, hipMemcpyDeviceToHost, g.s));
   if (grads) HIPC(hipMemcpyAsync(grads, N->G, N->nparam * 4, hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

int bsahip_mlp_forward(bsahip_mlp *N, const float *x, int32_t n, float *logits)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!N || !x || !logits || n < 1) return fail(BSAHIP_EINVAL, "bad MLP forward arguments");
   float *xd, *xn, *z;
   CHK(N->B.get("fx", (size_t)n * N->nin, &xd));
   HIPC(hipMemcpyAsync(xd, x, (size_t)n * N->nin * 4, hipMemcpyHostToDevice, g.s));
   CHK(mlp_normalise(N, xd, n, &xn, "fxn"));
   CHK(mlp_forward(N, xn, n, &z));
   HIPC(hipMemcpyAsync(logits, z, (size_t)n * 4, hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

int bsahip_mlp_loss_grad(bsahip_mlp *N, const float *x, const float *y, const float *w, int32_t n,
                         float *loss, float *grads)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!N || !x || !y || !w || n < 1) return fail(BSAHIP_EINVAL, "bad MLP gradient arguments");
   float *xd, *yd, *wd, *xn, *ld;
   CHK(N->B.get("gx", (size_t)n * N->nin, &xd));
   CHK(N->B.get("gy", (size_t)n, &yd));
   CHK(N->B.get("gw", (size_t)n, &wd));
   CHK(N->B.get("gl", 1, &ld));
   HIPC(hipMemcpyAsync(xd, x, (size_t)n * N->nin * 4, hipMemcpyHostToDevice, g.s));
   HIPC(hipMemcpyAsync(yd, y, (size_t)n * 4, hipMemcpyHostToDevice, g.s));
   HIPC(hipMemcpyAsync(wd, w, (size_t)n * 4, hipMemcpyHostToDevice, g.s));
   double ws = 0.0;
   for (int i = 0; i < n; ++i) ws += w[i];
   if (!(ws > 0.0)) return fail(BSAHIP_EINVAL, "sample weights must sum to > 0");
   CHK(mlp_normalise(N, xd, n, &xn, "gxn"));
   CHK(mlp_backward(N, xn, yd, wd, n, (float)ws, ld));
   if (loss) HIPC(hipMemcpyAsync(loss, ld, 4, hipMemcpyDeviceToHost, g.s));
   if (grads) HIPC(hipMemcpyAsync(grads, N->G, N->nparam * 4, hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

// Minibatch Adam over a resident data set: epochs passes, each over a seeded
// permutation in batches of `batch` rows.  losses[e] = mean batch loss of epoch e.
int bsahip_mlp_train(bsahip_mlp *N, const float *x, const float *y, const float *w, int32_t n,
                     int32_t epochs, int32_t batch, float lr, float weight_decay, uint64_t seed,
                     float *losses)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!N || !x || !y || !w || n < 1 || epochs < 1 || batch < 1)
      return fail(BSAHIP_EINVAL, "bad MLP training arguments");
   float *xd, *yd, *wd, *xn, *xb, *yb, *wb, *ld;
   int *idx;
   CHK(N->B.get("tx", (size_t)n * N->nin, &xd));
   CHK(N->B.get("ty", (size_t)n, &yd));
   CHK(N->B.get("tw", (size_t)n, &wd));
   HIPC(hipMemcpyAsync(xd, x, (size_t)n * N->nin * 4, hipMemcpyHostToDevice, g.s));
   HIPC(hipMemcpyAsync(yd, y, (size_t)n * 4, hipMemcpyHostToDevice, g.s));
   HIPC(hipMemcpyAsync(wd, w, (size_t)n * 4, hipMemcpyHostToDevice, g.s));
   CHK(mlp_normalise(N, xd, n, &xn, "txn"));
   const int bs = std::min(batch, (int)n);
   CHK(N->B.get("tbx", (size_t)bs * N->nin, &xb));
   CHK(N->B.get("tby", (size_t)bs, &yb));
   CHK(N->B.get("tbw", (size_t)bs, &wb));
   CHK(N->B.get("tidx", (size_t)bs, &idx));
   CHK(N->B.get("tl", 1, &ld));
   std::vector<int> perm(n);
   for (int i = 0; i < n; ++i) perm[i] = i;
   uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
   auto rnd = [&]() {
      st ^= st << 13;
      st ^= st >> 7;
      st ^= st << 17;
      return st;
   };
   const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
   for (int e = 0; e < epochs; ++e) {
      for (int i = n - 1; i > 0; --i) std::swap(perm[i], perm[rnd() % (uint64_t)(i + 1)]);
      double esum = 0.0;
      int nb = 0;
      for (int s0 = 0; s0 + bs <= n; s0 += bs) {
         HIPC(hipMemcpyAsync(idx, perm.data() + s0, (size_t)bs * 4, hipMemcpyHostToDevice, g.s));
         const long tot = (long)bs * N->nin;
         k_mlp_gather<<<(unsigned)((tot + 255) / 256), 256, 0, g.s>>>(xn, yd, wd, idx, bs, N->nin,
                                                                     xb, yb, wb);
         HIPC(hipGetLastError());
         double ws = 0.0;
         for (int r = 0; r < bs; ++r) ws += w[perm[s0 + r]];
         if (!(ws > 0.0)) continue;
         CHK(mlp_backward(N, xb, yb, wb, bs, (float)ws, ld));
         ++N->step;
         const float c1 = 1.f - std::pow(b1, (float)N->step), c2 = 1.f - std::pow(b2, (float)N->step);
         k_mlp_adam<<<(unsigned)((N->nparam + 255) / 256), 256, 0, g.s>>>(
            N->P, N->G, N->Mo, N->Vo, (long)N->nparam, lr, b1, b2, eps, c1, c2, weight_decay);
         HIPC(hipGetLastError());
         float lh = 0.f;
         HIPC(hipMemcpyAsync(&lh, ld, 4, hipMemcpyDeviceToHost, g.s));
         HIPC(hipStreamSynchronize(g.s));
         esum += lh;
         ++nb;
      }
      if (losses) losses[e] = nb ? (float)(esum / nb) : 0.f;
   }
   return 0;
}

int bsahip_mlp_destroy(bsahip_mlp *N)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   delete N;
   return 0;
}

}  // extern "C"
