"""ctypes binding of the BsaCL C ABI exported by libbsahip.so (include/bsacl.h).

Mirrors the reference's Fortran module BsaCL (BsaCL/BsaCL_Fortran/bsacl.F90:
54-205, shape-deriving wrappers bsacl_impl.F90:18-287): every ``acquire_*``
passes the caller's array by pointer and keeps a reference to it, as the C
runtime stores raw pointers until bsaclRun consumes them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .abi import lib


class BsaCL:
    def __init__(self):
        self.L = lib()
        self._keep = {}

    def _ptr(self, name, a, dtype, ct):
        arr = np.require(a, dtype=dtype, requirements=["F_CONTIGUOUS", "ALIGNED"])
        self._keep[name] = arr
        return arr.ctypes.data_as(C.POINTER(ct))

    # bsacl.F90 bsacl_Init / bsacl_InitDeviceMemory / bsacl_Run / bsacl_Finalise
    def init(self, n_threads: int = 1) -> int:
        return self.L.bsaclInit(C.c_uint(n_threads))

    def init_device_memory(self) -> int:
        return self.L.bsaclInitDeviceMemory()

    def run(self, dim: int, i_thread: int = 0, h_res=None):
        out = np.zeros(dim) if h_res is None else h_res
        rc = self.L.bsaclRun(C.c_uint(i_thread), C.c_size_t(dim),
                             out.ctypes.data_as(C.POINTER(C.c_double)))
        return rc, out

    def set_kernel_id(self, kid: int) -> int:
        return self.L.bsaclSetKernelID(C.c_uint(kid))

    def finalise(self):
        self.L.bsaclFinalise()
        self._keep.clear()

    def abort(self, ierr: int):
        self.L.bsaclAbort(C.c_int(ierr))

    def set_device_type(self, t: int):
        self.L.bsaclSetDeviceType(C.c_uint32(t))

    def verify_max_alloc(self, idim: int) -> int:
        ican = C.c_uint32(7)
        self.L.bsaclVerifyMaxAllocCondition(C.c_size_t(idim), C.byref(ican))
        return ican.value

    # acquire (bsacl_impl.F90 derives the sizes from the array shapes)
    def acquire_psd_id(self, pid: int):
        self.L.bsaclAcquirePSDId(C.c_uint32(pid))

    def acquire_struct_mod_mat(self, phi, natf):
        self.L.bsaclAcquireStructModMat(self._ptr("phi", phi, np.float64, C.c_double),
                                        self._ptr("natf", natf, np.float64, C.c_double),
                                        C.c_uint32(np.shape(phi)[0]), C.c_uint32(np.shape(phi)[1]))

    def acquire_modal_matrices(self, Mg, Cg, Kg):
        self.L.bsaclAcquireModalMatrices(self._ptr("Mg", Mg, np.float64, C.c_double),
                                         self._ptr("Cg", Cg, np.float64, C.c_double),
                                         self._ptr("Kg", Kg, np.float64, C.c_double))

    def acquire_loaded_nodes_list(self, nodes):
        self.L.bsaclAcquireLoadedNodesList(self._ptr("nl", nodes, np.int32, C.c_int),
                                           C.c_uint32(np.size(nodes)))

    def acquire_total_n_of_nodes(self, nn: int):
        self.L.bsaclAcquireTotalNOfNodes(C.c_uint32(nn))

    def acquire_used_modes_list(self, modes):
        self.L.bsaclAcquireUsedModesList(self._ptr("modes", modes, np.int32, C.c_int),
                                         C.c_uint32(np.size(modes)))

    def acquire_wind_coeffs(self, wfc):
        self.L.bsaclAcquireWindCoeffs(self._ptr("wfc", wfc, np.float64, C.c_double),
                                      C.c_uint32(np.shape(wfc)[2]), C.c_uint32(np.shape(wfc)[0]),
                                      C.c_uint32(np.shape(wfc)[1]))

    def acquire_phi_times_c_mat(self, psi):
        self.L.bsaclAcquirePhiTimesCMat(self._ptr("psi", psi, np.float64, C.c_double),
                                        C.c_uint32(psi.shape[0]), C.c_uint32(psi.shape[1]),
                                        C.c_uint32(psi.shape[2]))

    def acquire_turb_components_list(self, tc):
        self.L.bsaclAcquireTurbComponentsList(self._ptr("tc", tc, np.int32, C.c_int),
                                              C.c_uint32(np.size(tc)))

    def acquire_nodal_correlation(self, corr):
        self.L.bsaclAcquireNodalCorrelation(self._ptr("corr", corr, np.float64, C.c_double),
                                            C.c_uint32(np.shape(corr)[0]))

    def acquire_wind_nodal_velocities(self, u):
        self.L.bsaclAcquireWindNodalVelocities(self._ptr("u", u, np.float64, C.c_double))

    def acquire_wind_nodal_wind_zones(self, wz):
        self.L.bsaclAcquireWindNodalWindZones(self._ptr("wz", wz, np.int32, C.c_int))

    def acquire_wind_turb_scales(self, L):
        self.L.bsaclAcquireWindTurbScales(self._ptr("L", L, np.float64, C.c_double),
                                          C.c_uint32(np.shape(L)[2]))

    def acquire_wind_turb_std(self, s):
        self.L.bsaclAcquireWindTurbStd(self._ptr("s", s, np.float64, C.c_double),
                                       C.c_uint32(np.shape(s)[1]))

    def acquire_computation_freqs(self, fi, fj, i_thread: int = 0):
        self.L.bsaclAcquireComputationFreqs(C.c_uint(i_thread), C.c_uint32(np.size(fi)),
                                            self._ptr(f"fi{i_thread}", fi, np.float64, C.c_double),
                                            C.c_uint32(np.size(fj)),
                                            self._ptr(f"fj{i_thread}", fj, np.float64, C.c_double))

    def acquire_problem(self, prob, psi, modes):
        """The acquisition sequence of bsa_Run (src/bsa/BsaLibImpl.F90:269-289)."""
        mo = np.asarray(modes) - 1
        self.acquire_psd_id(prob.psd_id)
        self.acquire_struct_mod_mat(prob.phi, prob.natf)
        self.acquire_modal_matrices(prob.Mm[mo], prob.Cm[np.ix_(mo, mo)], prob.Km[mo])
        self.acquire_loaded_nodes_list(prob.n_load)
        self.acquire_total_n_of_nodes(prob.nn)
        self.acquire_used_modes_list(modes)
        self.acquire_wind_coeffs(prob.wfc)
        self.acquire_phi_times_c_mat(psi)
        self.acquire_turb_components_list(prob.tc)
        self.acquire_nodal_correlation(prob.nod_corr)
        self.acquire_wind_nodal_velocities(prob.u_node)
        self.acquire_wind_nodal_wind_zones(prob.wz_node)
        self.acquire_wind_turb_scales(prob.turb_scales)
        self.acquire_wind_turb_std(prob.turb_std)
