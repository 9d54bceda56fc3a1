"""Python host API over the C ABI of libbnsamd.so (include/bns_nurates_amd.h).

Mirrors the reference's host interface for the gray M1 path:

* ``gauss_legendre(n)``             -- ``GaussLegendre`` (integration.hpp:28-87)
* ``Context.compute_m1_opacities``  -- batched ``ComputeM1Opacities`` /
  ``ComputeM1OpacitiesNotStimulated`` (m1_opacities.hpp:1297-1313)
* ``Context.compute_spectral_opacities`` -- batched
  ``ComputeSpectralOpacitiesNotStimulatedAbs`` / ``...StimulatedAbs``
  (m1_opacities.hpp:1505-1660) over points x energies
* ``Context.distr_from_m1``         -- ``CalculateDistrParamsFromM1``
  (distribution.hpp:323-333)
* ``Context.equilibrium``           -- ``NuEquilibriumParams`` +
  ``ComputeM1DensitiesEq`` (distribution.hpp:253-290, 447-473)

Point data travel as SoA blocks (21 / 28 / 20 fields of npts float64, the
field order of the header).  Device blocks are torch tensors on the HIP
device (torch is used for device memory and streams only); host blocks are
numpy arrays.  The HIP library is required: there is no CPU fallback, and a
missing or unloadable library raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbnsamd.so")

PT_NFIELDS, DI_NFIELDS, OUT_NFIELDS = 21, 28, 20
SP_NFIELDS = 16  # SpectralOpacities: j, kappa, j_s, kappa_s per species
FP64, MIXED = 0, 1
BREM_HR98, BREM_BRT06, BREM_GP19 = 0, 1, 2

_STATUS = {0: "OK", 1: "invalid argument", 2: "HIP error", 3: "no HIP device",
           4: "no quadrature set", 5: "out of device memory"}


class BnsAmdError(RuntimeError):
    pass


class _Flags(C.Structure):
    _fields_ = [(n, C.c_int) for n in
                ("use_abs_em", "use_pair", "use_brem", "use_inelastic_scatt", "use_iso")]


class _Params(C.Structure):
    _fields_ = [(n, C.c_int) for n in
                ("use_dU", "use_dm_eff", "use_WM_ab", "use_WM_sc", "use_decay",
                 "brem_implementation", "use_NN_medium_corr", "neglect_blocking")]


@dataclass
class OpacityFlags:
    """``OpacityFlags`` (bns_nurates.hpp:418-438)."""
    use_abs_em: int = 1
    use_pair: int = 1
    use_brem: int = 1
    use_inelastic_scatt: int = 1
    use_iso: int = 1

    @classmethod
    def from_array(cls, a):
        return cls(*[int(v) for v in a])

    def c(self):
        return _Flags(self.use_abs_em, self.use_pair, self.use_brem, self.use_inelastic_scatt,
                      self.use_iso)


@dataclass
class OpacityParams:
    """``OpacityParams`` (bns_nurates.hpp:338-370); defaults are
    ``opacity_params_default_all`` with the bremsstrahlung model to choose."""
    use_dU: int = 1
    use_dm_eff: int = 1
    use_WM_ab: int = 1
    use_WM_sc: int = 1
    use_decay: int = 1
    brem_implementation: int = BREM_HR98
    use_NN_medium_corr: int = 1
    neglect_blocking: int = 1

    @classmethod
    def from_array(cls, a):
        return cls(*[int(v) for v in a])

    def c(self):
        return _Params(self.use_dU, self.use_dm_eff, self.use_WM_ab, self.use_WM_sc,
                       self.use_decay, self.brem_implementation, self.use_NN_medium_corr,
                       self.neglect_blocking)


def as_flags(flags) -> OpacityFlags:
    """OpacityFlags from None (all reactions on), a sequence of 5 ints or an
    OpacityFlags."""
    if flags is None:
        return OpacityFlags()
    if isinstance(flags, OpacityFlags):
        return flags
    a = [int(v) for v in np.asarray(flags).ravel()]
    if len(a) != 5:
        raise BnsAmdError(f"flags: expected 5 values, got {len(a)}")
    return OpacityFlags(*a)


def as_params(params) -> OpacityParams:
    """OpacityParams from None (default_all, HR98), a sequence of 8 ints or
    an OpacityParams."""
    if params is None:
        return OpacityParams()
    if isinstance(params, OpacityParams):
        return params
    a = [int(v) for v in np.asarray(params).ravel()]
    if len(a) != 8:
        raise BnsAmdError(f"params: expected 8 values, got {len(a)}")
    return OpacityParams(*a)


class _ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget_bytes", C.c_size_t), ("streams", C.c_int),
                ("max_batch_points", C.c_int64), ("chunk_points", C.c_int64),
                ("starvation_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "requests_submitted", "requests_completed", "points_completed", "batches",
        "max_inflight_bytes", "max_concurrent_batches", "arena_bytes")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


_lib = None


def load_library():
    """Load libbnsamd.so (build it first: ``make -C bns_nurates_amd/csrc``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BnsAmdError(f"{LIB_PATH} is not built: run make -C bns_nurates_amd/csrc")
    # One HIP runtime per process.  PyTorch's ROCm wheels bundle their own
    # libamdhip64 and load it by file name; loaded before torch, this
    # library would bring in the system copy (same soname) and torch would
    # then load a second runtime that cannot share the device with the
    # first.  Importing torch first makes this library bind to torch's copy.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    vp, i64, dp = C.c_void_p, C.c_int64, C.POINTER(C.c_double)
    sig = {
        "bnsamd_version": ([], C.c_int),
        "bnsamd_create": ([C.c_int, C.POINTER(vp)], C.c_int),
        "bnsamd_destroy": ([vp], None),
        "bnsamd_last_error": ([vp], C.c_char_p),
        "bnsamd_set_stream": ([vp, vp], C.c_int),
        "bnsamd_get_stream": ([vp], vp),
        "bnsamd_synchronize": ([vp], C.c_int),
        "bnsamd_gauss_legendre": ([C.c_int, C.c_double, C.c_double, dp, dp], C.c_int),
        "bnsamd_set_quadrature": ([vp, C.c_int, dp, dp, C.c_int, dp, dp], C.c_int),
        "bnsamd_set_gp19_table": ([vp, C.c_int, dp, dp, C.c_int, dp, C.c_int, dp, C.c_int, dp,
                                   dp, dp], C.c_int),
        "bnsamd_m1_opacities": ([vp, i64, i64, vp, vp, C.POINTER(_Flags), C.POINTER(_Params),
                                 C.c_int, C.c_int, vp], C.c_int),
        "bnsamd_m1_opacities_host": ([vp, i64, vp, vp, C.POINTER(_Flags), C.POINTER(_Params),
                                      C.c_int, C.c_int, vp], C.c_int),
        "bnsamd_m1_opacities_host_static": ([vp, i64, vp, vp, C.POINTER(_Flags),
                                             C.POINTER(_Params), C.c_int, C.c_int, vp], C.c_int),
        "bnsamd_spectral_opacities": ([vp, i64, i64, vp, vp, i64, vp, i64, C.POINTER(_Flags),
                                       C.POINTER(_Params), C.c_int, vp], C.c_int),
        "bnsamd_spectral_opacities_aos": ([vp, i64, i64, vp, vp, i64, vp, i64,
                                           C.POINTER(_Flags), C.POINTER(_Params), C.c_int, vp],
                                          C.c_int),
        "bnsamd_spectral_opacities_host": ([vp, i64, vp, vp, i64, vp, C.c_int,
                                            C.POINTER(_Flags), C.POINTER(_Params), C.c_int,
                                            vp], C.c_int),
        "bnsamd_distr_from_m1": ([vp, i64, i64, vp, vp], C.c_int),
        "bnsamd_equilibrium": ([vp, i64, i64, vp, vp], C.c_int),
        "bnsamd_m1_densities_eq": ([vp, i64, i64, vp, vp], C.c_int),
        "bnsamd_distr_from_m1_host": ([vp, i64, vp, vp], C.c_int),
        "bnsamd_equilibrium_host": ([vp, i64, vp, vp], C.c_int),
        "bnsamd_m1_densities_eq_host": ([vp, i64, vp, vp], C.c_int),
        "bnsamd_fdi": ([vp, C.c_int, i64, vp, vp], C.c_int),
        "bnsamd_server_create": ([C.POINTER(_ServerConfig), C.POINTER(vp)], C.c_int),
        "bnsamd_server_destroy": ([vp], None),
        "bnsamd_server_last_error": ([vp], C.c_char_p),
        "bnsamd_server_submit": ([vp, i64, vp, C.c_int, C.POINTER(_Flags), C.POINTER(_Params),
                                  C.c_int, C.c_int, vp, C.POINTER(i64)], C.c_int),
        "bnsamd_server_poll": ([vp, i64, C.POINTER(C.c_int)], C.c_int),
        "bnsamd_server_wait": ([vp, i64, C.POINTER(C.c_double)], C.c_int),
        "bnsamd_server_get_stats": ([vp, C.POINTER(ServerStats)], C.c_int),
        "bnsamd_host_alloc": ([C.c_size_t, C.POINTER(vp)], C.c_int),
        "bnsamd_host_free": ([vp], None),
        "bnsamd_mlp_create": ([C.c_int, C.c_int, C.c_int, C.POINTER(vp)], C.c_int),
        "bnsamd_mlp_destroy": ([vp], None),
        "bnsamd_mlp_last_error": ([vp], C.c_char_p),
        "bnsamd_mlp_param_count": ([], C.c_int),
        "bnsamd_mlp_set_params": ([vp, C.c_int, vp], C.c_int),
        "bnsamd_mlp_get_params": ([vp, C.c_int, vp], C.c_int),
        "bnsamd_mlp_set_norm": ([vp, vp], C.c_int),
        "bnsamd_mlp_gradients": ([vp, C.c_int, C.c_int, vp, vp, vp, vp], C.c_int),
        "bnsamd_mlp_train": ([vp, C.c_int, i64, vp, vp, vp, C.c_int, C.c_int, C.c_float,
                              C.c_float, vp], C.c_int),
        "bnsamd_mlp_train_ensemble": ([vp, i64, vp, vp, vp, C.c_int, C.c_int, C.c_float,
                                       C.c_float, vp], C.c_int),
        "bnsamd_mlp_predict": ([vp, C.c_int, C.c_int, vp, vp], C.c_int),
        "bnsamd_mlp_set_domain": ([vp, vp], C.c_int),
        "bnsamd_mlp_set_acceptance": ([vp, vp, vp], C.c_int),
        "bnsamd_mlp_set_judge": ([vp, vp, C.c_double], C.c_int),
        "bnsamd_mlp_infer": ([vp, vp, i64, i64, vp, vp, vp, vp], C.c_int),
        "bnsamd_surrogate_m1_opacities": ([vp, vp, i64, i64, vp, C.c_float, C.POINTER(_Flags),
                                           C.POINTER(_Params), C.c_int, vp, C.POINTER(i64)],
                                          C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def gauss_legendre(n: int, x1: float = 0.0, x2: float = 1.0):
    """Nodes and weights of ``GaussLegendre`` on [x1, x2] (host)."""
    lib = load_library()
    x = np.zeros(n)
    w = np.zeros(n)
    rc = lib.bnsamd_gauss_legendre(int(n), float(x1), float(x2), _dptr(x), _dptr(w))
    if rc:
        raise BnsAmdError(f"bnsamd_gauss_legendre: {_STATUS.get(rc, rc)}")
    return x, w


def _tensor_ptr(t, nfields: int, npts: int, name: str):
    """Device pointer and leading dimension of a [nfields, >= npts] float64
    device tensor with unit stride along points."""
    if t.dtype.__str__() != "torch.float64":
        raise BnsAmdError(f"{name}: float64 tensor required, got {t.dtype}")
    if not t.is_cuda:
        raise BnsAmdError(f"{name}: device tensor required")
    if t.dim() != 2 or t.shape[0] != nfields or t.shape[1] < npts or t.stride(1) != 1:
        raise BnsAmdError(f"{name}: expected shape [{nfields}, >= {npts}] with unit point stride")
    return t.data_ptr(), t.stride(0)


class Context:
    """A device context: stream, resident quadrature tables, staging buffers."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.bnsamd_create(int(device), C.byref(h))
        if rc:
            raise BnsAmdError(f"bnsamd_create(device={device}): {_STATUS.get(rc, rc)}")
        self.h = h
        self.device = device
        self._quad = None
        self._stream_explicit = False

    def close(self):
        if getattr(self, "h", None):
            self.lib.bnsamd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            msg = self.lib.bnsamd_last_error(self.h).decode(errors="replace")
            raise BnsAmdError(f"{what}: {_STATUS.get(rc, rc)}: {msg}")

    def set_stream(self, stream_handle):
        """Launch on a HIP stream (an int handle such as
        torch.cuda.current_stream().cuda_stream; 0 is HIP's default stream).
        None returns to following torch's current stream at every call."""
        if stream_handle is not None:
            self._check(self.lib.bnsamd_set_stream(self.h, C.c_void_p(stream_handle)),
                        "bnsamd_set_stream")
        self._stream_explicit = stream_handle is not None

    def _follow_torch_stream(self):
        """Device tensors are produced on torch's current stream: unless a
        stream was set explicitly, launch there so that no copy is raced."""
        if not self._stream_explicit:
            import torch
            s = torch.cuda.current_stream(self.device).cuda_stream
            self._check(self.lib.bnsamd_set_stream(self.h, C.c_void_p(s)), "bnsamd_set_stream")

    @property
    def stream(self) -> int:
        return self.lib.bnsamd_get_stream(self.h) or 0

    def synchronize(self):
        self._check(self.lib.bnsamd_synchronize(self.h), "bnsamd_synchronize")

    def set_quadrature(self, n2: int, n1: int | None = None, quad_2d=None, quad_1d=None):
        """Select the quadratures: orders (with the library's GaussLegendre on
        [0, 1]) or explicit (x, w) pairs."""
        n1 = n2 if n1 is None else n1
        x2, w2 = quad_2d if quad_2d is not None else gauss_legendre(n2)
        x1, w1 = quad_1d if quad_1d is not None else gauss_legendre(n1)
        x1, w1, x2, w2 = (np.ascontiguousarray(a, dtype=np.float64) for a in (x1, w1, x2, w2))
        self._check(self.lib.bnsamd_set_quadrature(self.h, len(x1), _dptr(x1), _dptr(w1),
                                                   len(x2), _dptr(x2), _dptr(w2)),
                    "bnsamd_set_quadrature")
        self._quad = (len(x1), len(x2))

    def set_gp19_table(self, table):
        """Installs the GP19 bremsstrahlung table (a dict with the axes nb
        [fm^-3], lognb, ye, T [MeV], w [MeV], logw and data [nnb, nye, nT,
        nw]; bnsamd_set_gp19_table) or clears it (None)."""
        if table is None:
            self._check(self.lib.bnsamd_set_gp19_table(self.h, 0, None, None, 0, None, 0, None, 0,
                                                       None, None, None), "bnsamd_set_gp19_table")
            return
        ax = {k: np.ascontiguousarray(table[k], dtype=np.float64)
              for k in ("nb", "lognb", "ye", "T", "w", "logw")}
        data = np.ascontiguousarray(table["data"], dtype=np.float64)
        if data.shape != (len(ax["nb"]), len(ax["ye"]), len(ax["T"]), len(ax["w"])):
            raise BnsAmdError("gp19 table: data must be [nnb, nye, nT, nw]")
        self._check(self.lib.bnsamd_set_gp19_table(
            self.h, len(ax["nb"]), _dptr(ax["nb"]), _dptr(ax["lognb"]), len(ax["ye"]),
            _dptr(ax["ye"]), len(ax["T"]), _dptr(ax["T"]), len(ax["w"]), _dptr(ax["w"]),
            _dptr(ax["logw"]), _dptr(data)), "bnsamd_set_gp19_table")

    def compute_m1_opacities(self, pts, flags=None, params=None, stim_abs: int = 1,
                             precision: int = FP64, distr=None, out=None, npts=None):
        """Gray M1 opacities [20, npts] of the point block ``pts`` [21, npts].

        Device tensors run asynchronously on the context stream and write
        ``out`` (allocated when None); numpy arrays go through the
        synchronous host entry point and return a numpy array (``out``, a
        C-contiguous float64 [20, npts] array such as a PinnedBuffer's,
        when given)."""
        fl, pr = as_flags(flags).c(), as_params(params).c()
        if isinstance(pts, np.ndarray):
            pts = np.ascontiguousarray(pts, dtype=np.float64)
            n = pts.shape[1] if npts is None else npts
            if pts.shape[0] != PT_NFIELDS or pts.shape[1] != n:
                raise BnsAmdError(f"pts: expected a [{PT_NFIELDS}, {n}] host array")
            if out is None:
                res = np.empty((OUT_NFIELDS, n))
            elif (isinstance(out, np.ndarray) and out.dtype == np.float64 and
                  out.flags.c_contiguous and out.shape == (OUT_NFIELDS, n)):
                res = out
            else:
                raise BnsAmdError(f"out: expected a C-contiguous float64 [{OUT_NFIELDS}, {n}] "
                                  "host array")
            dd = None
            if distr is not None:
                distr = np.ascontiguousarray(distr, dtype=np.float64)
                dd = distr.ctypes.data
            self._check(self.lib.bnsamd_m1_opacities_host(
                self.h, n, pts.ctypes.data, dd, C.byref(fl), C.byref(pr), int(stim_abs),
                int(precision), res.ctypes.data), "bnsamd_m1_opacities_host")
            return res
        import torch
        self._follow_torch_stream()
        n = pts.shape[1] if npts is None else npts
        pp, ld = _tensor_ptr(pts, PT_NFIELDS, n, "pts")
        dp = None
        if distr is not None:
            dp, ldd = _tensor_ptr(distr, DI_NFIELDS, n, "distr")
            if ldd != ld:
                raise BnsAmdError("distr must have the leading dimension of pts")
        if out is None:
            out = torch.empty((OUT_NFIELDS, ld), dtype=torch.float64, device=pts.device)
        op, ldo = _tensor_ptr(out, OUT_NFIELDS, n, "out")
        if ldo != ld:
            raise BnsAmdError("out must have the leading dimension of pts")
        self._check(self.lib.bnsamd_m1_opacities(self.h, n, ld, pp, dp, C.byref(fl), C.byref(pr),
                                                 int(stim_abs), int(precision), op),
                    "bnsamd_m1_opacities")
        return out

    def compute_spectral_opacities(self, pts, nu, flags=None, params=None, stim_abs: int = 0,
                                   distr=None, out=None, layout: str = "fields"):
        """Spectral opacities [16, nnu, npts] (j, kappa, j_s, kappa_s of the
        four species; field order of SpectralOpacities) of the point block
        ``pts`` [21, npts] at the energies ``nu``: [nnu] shared by all points
        or [nnu, npts] per point.  The partner-energy integrals use the 1D
        quadrature of ``set_quadrature``.  stim_abs = 0 is
        ComputeSpectralOpacitiesNotStimulatedAbs, 1 ...StimulatedAbs.

        Device tensors run asynchronously on the context stream (``out`` a
        [16 * nnu, ld] tensor, allocated when None, returned viewed as
        [16, nnu, ld]); numpy arrays go through the synchronous host entry
        point.  ``layout="struct"`` (device tensors): one SpectralOpacities
        record per (point, energy), ``out`` a contiguous [npts, nnu, 16]
        tensor (bnsamd_spectral_opacities_aos)."""
        if layout not in ("fields", "struct"):
            raise BnsAmdError("layout must be 'fields' or 'struct'")
        fl, pr = as_flags(flags).c(), as_params(params).c()
        if isinstance(pts, np.ndarray):
            if layout != "fields":
                raise BnsAmdError("layout='struct' takes device tensors")
            pts = np.ascontiguousarray(pts, dtype=np.float64)
            n = pts.shape[1]
            nu = np.ascontiguousarray(nu, dtype=np.float64)
            if nu.ndim == 1:
                per_point = 0
            elif nu.ndim == 2 and nu.shape[1] == n:
                per_point = 1
            else:
                raise BnsAmdError("nu must be [nnu] or [nnu, npts]")
            nnu = nu.shape[0]
            res = np.empty((SP_NFIELDS, nnu, n))
            dd = None
            if distr is not None:
                distr = np.ascontiguousarray(distr, dtype=np.float64)
                dd = distr.ctypes.data
            self._check(self.lib.bnsamd_spectral_opacities_host(
                self.h, n, pts.ctypes.data, dd, nnu, nu.ctypes.data, per_point, C.byref(fl),
                C.byref(pr), int(stim_abs), res.ctypes.data), "bnsamd_spectral_opacities_host")
            return res
        import torch
        self._follow_torch_stream()
        n = pts.shape[1]
        pp, ld = _tensor_ptr(pts, PT_NFIELDS, n, "pts")
        if nu.dtype != torch.float64 or not nu.is_cuda or not nu.is_contiguous():
            raise BnsAmdError("nu: contiguous float64 device tensor required")
        if nu.dim() == 1:
            nu_ld = 0
        elif nu.dim() == 2 and nu.shape[1] >= n:
            nu_ld = nu.stride(0)
        else:
            raise BnsAmdError("nu must be [nnu] or [nnu, >= npts]")
        nnu = nu.shape[0]
        dp = None
        if distr is not None:
            dp, ldd = _tensor_ptr(distr, DI_NFIELDS, n, "distr")
            if ldd != ld:
                raise BnsAmdError("distr must have the leading dimension of pts")
        if layout == "struct":
            if out is None:
                out = torch.empty((n, nnu, SP_NFIELDS), dtype=torch.float64, device=pts.device)
            if (out.dtype != torch.float64 or not out.is_cuda or not out.is_contiguous()
                    or tuple(out.shape) != (n, nnu, SP_NFIELDS)):
                raise BnsAmdError(f"out: contiguous float64 device tensor [{n}, {nnu}, "
                                  f"{SP_NFIELDS}] required")
            self._check(self.lib.bnsamd_spectral_opacities_aos(
                self.h, n, ld, pp, dp, nnu, nu.data_ptr(), nu_ld, C.byref(fl), C.byref(pr),
                int(stim_abs), out.data_ptr()), "bnsamd_spectral_opacities_aos")
            return out
        if out is None:
            out = torch.empty((SP_NFIELDS * nnu, ld), dtype=torch.float64, device=pts.device)
        op, ldo = _tensor_ptr(out, SP_NFIELDS * nnu, n, "out")
        if ldo != ld:
            raise BnsAmdError("out must have the leading dimension of pts")
        self._check(self.lib.bnsamd_spectral_opacities(
            self.h, n, ld, pp, dp, nnu, nu.data_ptr(), nu_ld, C.byref(fl), C.byref(pr),
            int(stim_abs), op), "bnsamd_spectral_opacities")
        return out.view(SP_NFIELDS, nnu, ld)

    def distr_from_m1(self, pts, distr=None):
        import torch
        self._follow_torch_stream()
        n = pts.shape[1]
        pp, ld = _tensor_ptr(pts, PT_NFIELDS, n, "pts")
        if distr is None:
            distr = torch.empty((DI_NFIELDS, ld), dtype=torch.float64, device=pts.device)
        dp, ldd = _tensor_ptr(distr, DI_NFIELDS, n, "distr")
        if ldd != ld:
            # the kernel addresses both blocks with one leading dimension
            raise BnsAmdError("distr must have the leading dimension of pts")
        self._check(self.lib.bnsamd_distr_from_m1(self.h, n, ld, pp, dp), "bnsamd_distr_from_m1")
        return distr

    def equilibrium(self, pts, distr=None):
        """Writes equilibrium n, J into pts (in place) and returns the
        equilibrium distribution block."""
        import torch
        n = pts.shape[1]
        pp, ld = _tensor_ptr(pts, PT_NFIELDS, n, "pts")
        if distr is None:
            distr = torch.empty((DI_NFIELDS, ld), dtype=torch.float64, device=pts.device)
        dp, ldd = _tensor_ptr(distr, DI_NFIELDS, n, "distr")
        if ldd != ld:
            raise BnsAmdError("distr must have the leading dimension of pts")
        self._follow_torch_stream()
        self._check(self.lib.bnsamd_equilibrium(self.h, n, ld, pp, dp), "bnsamd_equilibrium")
        return distr

    def fdi(self, k: int, y):
        import torch
        self._follow_torch_stream()
        out = torch.empty_like(y)
        self._check(self.lib.bnsamd_fdi(self.h, int(k), y.numel(), y.data_ptr(), out.data_ptr()),
                    "bnsamd_fdi")
        return out


class PinnedBuffer:
    """Page-locked host memory (bnsamd_host_alloc) viewed as a float64 numpy
    array of the given shape; freed by close() or when collected."""

    def __init__(self, shape):
        self.lib = load_library()
        n = int(np.prod(shape))
        p = C.c_void_p()
        rc = self.lib.bnsamd_host_alloc(C.c_size_t(max(n, 1) * 8), C.byref(p))
        if rc:
            raise BnsAmdError(f"bnsamd_host_alloc({n * 8} bytes): {_STATUS.get(rc, rc)}")
        self.ptr = p
        self.array = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(max(n, 1),))
        self.array = self.array[:n].reshape(shape)

    def close(self):
        if getattr(self, "ptr", None):
            self.array = None
            self.lib.bnsamd_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request server (bnsamd_server_*): submit host SoA
    point blocks [21, npts] from any thread, wait for their [20, npts]
    outputs by id.  Requests above chunk_points points run as pieces that
    interleave with other requests; the smallest request is served first
    unless a piece has waited longer than starvation_ms."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 8 << 30, streams: int = 4,
                 max_batch_points: int = 1 << 20, chunk_points: int = 32768,
                 starvation_ms: float = 200.0):
        self.lib = load_library()
        cfg = _ServerConfig(int(device), int(hbm_budget_bytes), int(streams), int(max_batch_points),
                            int(chunk_points), float(starvation_ms))
        h = C.c_void_p()
        rc = self.lib.bnsamd_server_create(C.byref(cfg), C.byref(h))
        if rc:
            raise BnsAmdError(f"bnsamd_server_create: {_STATUS.get(rc, rc)}")
        self.h = h
        self._keep = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.bnsamd_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, pts: np.ndarray, out: np.ndarray, nquad: int, flags=None, params=None,
               stim_abs: int = 1, precision: int = FP64) -> int:
        """Queue a request; pts [21, npts] and out [20, npts] C-contiguous
        float64 (pinned buffers let the copies overlap).  Both must stay
        alive until wait() returns for the id."""
        if pts.dtype != np.float64 or not pts.flags.c_contiguous or pts.shape[0] != 21:
            raise BnsAmdError("submit: pts must be a C-contiguous float64 [21, npts] array")
        npts = pts.shape[1]
        if out.dtype != np.float64 or not out.flags.c_contiguous or out.shape != (20, npts):
            raise BnsAmdError("submit: out must be a C-contiguous float64 [20, npts] array")
        f, p = as_flags(flags).c(), as_params(params).c()
        rid = C.c_int64()
        rc = self.lib.bnsamd_server_submit(self.h, npts, pts.ctypes.data, int(nquad), C.byref(f),
                                           C.byref(p), int(stim_abs), int(precision),
                                           out.ctypes.data, C.byref(rid))
        if rc:
            msg = self.lib.bnsamd_server_last_error(self.h).decode(errors="replace")
            raise BnsAmdError(f"bnsamd_server_submit: {_STATUS.get(rc, rc)}: {msg}")
        self._keep[rid.value] = (pts, out)
        return rid.value

    def poll(self, rid: int) -> bool:
        d = C.c_int()
        rc = self.lib.bnsamd_server_poll(self.h, int(rid), C.byref(d))
        if rc:
            raise BnsAmdError(f"bnsamd_server_poll: {_STATUS.get(rc, rc)}")
        return bool(d.value)

    def wait(self, rid: int) -> float:
        """Blocks until the request is done; returns its latency in seconds."""
        lat = C.c_double()
        rc = self.lib.bnsamd_server_wait(self.h, int(rid), C.byref(lat))
        self._keep.pop(int(rid), None)
        if rc:
            raise BnsAmdError(f"bnsamd_server_wait({rid}): {_STATUS.get(rc, rc)}")
        return lat.value * 1e-6

    def stats(self) -> dict:
        st = ServerStats()
        self.lib.bnsamd_server_get_stats(self.h, C.byref(st))
        return st.as_dict()
