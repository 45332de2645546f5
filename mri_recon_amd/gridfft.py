"""GridFFT: the NUFFT/SENSE encoding operator on the HIP library.

Mirrors the reference's gridFFT (src/gridFFT.h:28-194) for torch tensors on a
HIP device: ``backward`` is A (image -> k-space, gridFFT::backward,
gridFFT.cpp:824-844), ``backward_residual`` is A X - d (856-876), ``forward``
is A^H (k-space -> image, gridFFT::forward, 791-812) and accumulates into X
like the reference.  Unlike the reference, one call covers every coil of a
frame (the caller's coil loop, recon_lib.cpp:1565-1567 and 1614-1622, moves
into the library).

Shapes: image (Nz, Ny, Nx) complex64; coil maps (C, Nz, Ny, Nx); k-space of a
frame (C, M); trajectory kx, ky, kz, kw float32 (M,) in the reference's units
(cycles/FOV, |k| <= N/2).
"""
import ctypes
import os

import torch

from . import _ffi
from ._ffi import MrnParams, check, lib


# the autotuner's choices, keyed by geometry, sample and coil counts and device
DEFAULT_TUNE_CACHE = os.environ.get(
    "MRN_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "mri_recon_amd", "tune.json"))


def _ptr(t):
    if t is None:
        return None
    if t.is_conj() or t.is_neg():
        # a lazily conjugated/negated view: its memory does not hold its values
        raise ValueError("pass a tensor with resolved conjugation (t.resolve_conj().resolve_neg())")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _need(t, dtype, shape=None, name="tensor"):
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor" % name)
    if t.device.type != "cuda":
        raise ValueError("%s must be on a HIP device" % name)
    if t.dtype != dtype:
        raise TypeError("%s must be %s, got %s" % (name, dtype, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError("%s must have shape %s, got %s" % (name, tuple(shape), tuple(t.shape)))
    return t


class _DevArray:
    """A device allocation owned by the op, exposed to torch through the
    array interface torch reads for HIP memory (__cuda_array_interface__)."""

    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<c8", "data": (ptr, False), "version": 3,
                                         "strides": None}


def _device_buffer(ptr, n):
    return torch.as_tensor(_DevArray(ptr, n), device="cuda")


def _trip(v):
    return (v, v, v) if not isinstance(v, (tuple, list)) else tuple(v)


class GridFFT:
    """One operator instance: geometry, tables, windows, coil maps, per-frame
    trajectories (and measured data for the solvers)."""

    def __init__(self, N, ncoils=1, nframes=1, overgrid=1.5, dwin=-1.0, kernel="kaiser",
                 grid_in=(1, 1, 1), fft_in=(1, 1, 1), grid=-1.0, grid_scale=1.0, k_rad=9999.0,
                 kernel_mode="table", form="batched", coil_chunk=0, tile=0, max_visits=0,
                 bin_edge=0, xcd_run=0, xcd_run_gather=0, gather_tiles=False, scatter_valu=False, fft_stockham=False, xcd_run_fft=0,
                 gather_coils=0, fuse_frames=-1, args=None):
        L = lib()
        p = MrnParams()
        L.mrn_params_default(ctypes.byref(p))
        N = _trip(N)
        for a in range(3):
            p.N[a] = int(N[a])
            p.grid[a] = float(_trip(grid)[a])
            p.dwin[a] = float(_trip(dwin)[a])
            p.grid_in[a] = int(grid_in[a])
            p.fft_in[a] = int(fft_in[a])
            p.grid_scale[a] = float(_trip(grid_scale)[a])
            p.tile[a] = int(_trip(tile)[a])
        p.overgrid = float(overgrid)
        p.kernel = _ffi.KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
        p.k_rad = float(k_rad)
        p.ncoils = int(ncoils)
        p.nframes = int(nframes)
        p.kernel_mode = _ffi.KMODES[kernel_mode] if isinstance(kernel_mode, str) else int(kernel_mode)
        p.form = _ffi.FORMS[form] if isinstance(form, str) else int(form)
        p.coil_chunk = int(coil_chunk)
        p.max_visits = int(max_visits)
        p.bin_edge = int(bin_edge)
        p.xcd_run = int(xcd_run)
        p.xcd_run_gather = int(xcd_run_gather)
        p.gather_tiles = 2 if gather_tiles == 2 else int(bool(gather_tiles))
        p.scatter_valu = int(bool(scatter_valu))
        p.fft_stockham = int(bool(fft_stockham))
        p.xcd_run_fft = int(xcd_run_fft)
        p.gather_coils = int(gather_coils)
        p.fuse_frames = int(fuse_frames)
        if args:  # gridFFT::read_commandline flags
            arr = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
            check(L.mrn_parse_args(ctypes.byref(p), len(args), arr))
        h = ctypes.c_void_p()
        check(L.mrn_op_create(ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        self.params = p
        self.N = (int(p.N[0]), int(p.N[1]), int(p.N[2]))
        self.C = int(ncoils)
        self.F = int(nframes)
        self.form = form
        self.M = [0] * self.F
        g = self.geometry()
        self.S, self.og, self.beta = g["S"], g["og"], g["beta"]

    # ---------------------------------------------------------------- life
    def close(self):
        if getattr(self, "_h", None):
            lib().mrn_op_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, s):
        check(s, self._h)

    @property
    def image_shape(self):
        return (self.N[2], self.N[1], self.N[0])

    # ------------------------------------------------------------ geometry
    def geometry(self):
        S = (ctypes.c_int32 * 3)()
        og = (ctypes.c_int32 * 3)()
        beta = (ctypes.c_float * 3)()
        self._c(lib().mrn_get_geometry(self._h, S, og, beta))
        return dict(S=tuple(S), og=tuple(og), beta=tuple(beta))

    def _host_vec(self, fn, axis):
        n = ctypes.c_int()
        self._c(fn(self._h, axis, None, 0, ctypes.byref(n)))
        out = (ctypes.c_float * n.value)()
        self._c(fn(self._h, axis, out, n.value, ctypes.byref(n)))
        return torch.tensor(list(out), dtype=torch.float32)

    def table(self, axis=0):
        return self._host_vec(lib().mrn_get_table, axis)

    def window(self, axis=0):
        return self._host_vec(lib().mrn_get_window, axis)

    # ---------------------------------------------------------------- setup
    def set_traj(self, kx, ky, kz, kw, frame=0):
        M = kx.numel()
        for t, nm in ((kx, "kx"), (ky, "ky"), (kz, "kz"), (kw, "kw")):
            _need(t, torch.float32, (M,), nm)
        self._c(lib().mrn_set_traj(self._h, frame, M, _ptr(kx), _ptr(ky), _ptr(kz), _ptr(kw), _stream()))
        self.M[frame] = M

    def set_smaps(self, smaps):
        _need(smaps, torch.complex64, (self.C,) + self.image_shape, "smaps")
        self._c(lib().mrn_set_smaps(self._h, _ptr(smaps), _stream()))

    def set_kdata(self, data, frame=0):
        _need(data, torch.complex64, (self.C, self.M[frame]), "data")
        self._c(lib().mrn_set_kdata(self._h, frame, _ptr(data), _stream()))

    def autotune(self, frame=0, cache=DEFAULT_TUNE_CACHE):
        """Time the launch shapes on this frame and keep the fastest; the choice
        is cached in the JSON file `cache` (None: no cache).  Returns tuning()."""
        path = None
        if cache:
            os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
            path = os.fsencode(cache)
        self._c(lib().mrn_autotune(self._h, frame, path, _stream()))
        return self.tuning()

    def support(self, frame=0):
        """k-space support of a frame: (xrng (Sz, Sy, 2), yrng (Sz, 2)) int32 numpy
        arrays of the first and last x cell per row and row y per plane."""
        import numpy as np
        Sx, Sy, Sz = self.S
        x = np.zeros((Sz, Sy, 2), np.int32)
        y = np.zeros((Sz, 2), np.int32)
        self._c(lib().mrn_get_support(self._h, frame, ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(y.ctypes.data)))
        return x, y

    def frame_stats(self, frame=0):
        """Work lists of one frame's gridding (mrn_frame_stats as a dict)."""
        t = _ffi.MrnFrameStats()
        self._c(lib().mrn_get_frame_stats(self._h, frame, ctypes.byref(t)))
        return {name: getattr(t, name) for name, _ in t._fields_ if name != "reserved"}

    def tuning(self):
        """Launch shapes in use (mrn_tuning as a dict)."""
        t = _ffi.MrnTuning()
        self._c(lib().mrn_get_tuning(self._h, ctypes.byref(t)))
        out = {name: getattr(t, name) for name, _ in t._fields_ if name not in ("reserved", "tile")}
        out["tile"] = tuple(t.tile)
        return out

    # ------------------------------------------------------------- operator
    def backward(self, X, data=None, frame=0, sub=None):
        """A: data[c] = A_c X (minus sub when given); returns data (C, M)."""
        _need(X, torch.complex64, self.image_shape, "X")
        if data is None:
            data = torch.empty((self.C, self.M[frame]), dtype=torch.complex64, device=X.device)
        _need(data, torch.complex64, (self.C, self.M[frame]), "data")
        if sub is not None:
            _need(sub, torch.complex64, (self.C, self.M[frame]), "sub")
        self._c(lib().mrn_image_to_kspace(self._h, frame, _ptr(X), _ptr(data), _ptr(sub), _stream()))
        return data

    def backward_residual(self, X, sub, data=None, frame=0):
        return self.backward(X, data=data, frame=frame, sub=sub)

    def forward(self, data, X=None, frame=0, accumulate=True):
        """A^H: X += sum_c A_c^H (kw taper data_c); a new X starts at zero."""
        _need(data, torch.complex64, (self.C, self.M[frame]), "data")
        if X is None:
            X = torch.zeros(self.image_shape, dtype=torch.complex64, device=data.device)
        _need(X, torch.complex64, self.image_shape, "X")
        self._c(lib().mrn_kspace_to_image(self._h, frame, _ptr(data), _ptr(X), int(bool(accumulate)), _stream()))
        return X

    image_to_kspace = backward
    kspace_to_image = forward

    def normal(self, X, out=None, frame=0):
        """E'E X = sum_c A_c^H W A_c X."""
        _need(X, torch.complex64, self.image_shape, "X")
        if out is None:
            out = torch.empty_like(X)
        self._c(lib().mrn_normal(self._h, frame, _ptr(X), _ptr(out), _stream()))
        return out

    def coil_backward(self, X, smap, traj, sub=None, data=None):
        """Single-coil A with the trajectory given per call (plain form)."""
        kx, ky, kz, kw = traj
        M = kx.numel()
        if data is None:
            data = torch.empty(M, dtype=torch.complex64, device=X.device)
        self._c(lib().mrn_coil_backward(self._h, _ptr(X), _ptr(smap), _ptr(data), M, _ptr(kx), _ptr(ky),
                                        _ptr(kz), _ptr(kw), _ptr(sub), _stream()))
        return data

    def coil_forward(self, X, smap, data, traj):
        """Single-coil X += A^H (kw taper data) (plain form)."""
        kx, ky, kz, kw = traj
        self._c(lib().mrn_coil_forward(self._h, _ptr(X), _ptr(smap), _ptr(data), kx.numel(), _ptr(kx),
                                       _ptr(ky), _ptr(kz), _ptr(kw), _stream()))
        return X

    def coil_forward_sos(self, X, data, traj):
        """X += |w * img|^2 of one coil's data, no coil map (gridFFT::forward_sos)."""
        kx, ky, kz, kw = traj
        self._c(lib().mrn_coil_forward_sos(self._h, _ptr(X), _ptr(data), kx.numel(), _ptr(kx), _ptr(ky),
                                           _ptr(kz), _ptr(kw), _stream()))
        return X

    def set_k_rad(self, k_rad):
        """Change the A^H taper radius (the reference's public member k_rad)."""
        self._c(lib().mrn_set_k_rad(self._h, float(k_rad), _stream()))

    # -------------------------------------------------------------- solvers
    def cg(self, iters, X=None, graph=True, trace=True):
        """SENSE-CG over all frames (recon_lib.cpp:1523-1759).  Returns
        (X (F, Nz, Ny, Nx), trace (iters, 4) [sum|R|^2, Re a, Im a, sum|Rnew|^2])."""
        dev = torch.device("cuda")
        if X is None:
            X = torch.empty((self.F,) + self.image_shape, dtype=torch.complex64, device=dev)
        tr = (ctypes.c_double * (4 * max(iters, 1)))() if trace else None
        self._c(lib().mrn_cg(self._h, _ptr(X), int(iters), int(bool(graph)), tr, _stream()))
        tt = torch.tensor(list(tr), dtype=torch.float64).reshape(-1, 4)[:iters] if trace else None
        return X, tt

    def cg_init(self, X=None):
        """X = 0, R = P = E^H W d; returns X (F, Nz, Ny, Nx)."""
        if X is None:
            X = torch.empty((self.F,) + self.image_shape, dtype=torch.complex64, device="cuda")
        self._c(lib().mrn_cg_init(self._h, _ptr(X), _stream()))
        return X

    def cg_iterate(self, X, iters, graph=True, trace=True):
        """Continue CG for `iters` steps; returns the trace (iters, 4)."""
        tr = (ctypes.c_double * (4 * max(iters, 1)))() if trace else None
        self._c(lib().mrn_cg_iterate(self._h, _ptr(X), int(iters), int(bool(graph)), tr, _stream()))
        return torch.tensor(list(tr), dtype=torch.float64).reshape(-1, 4)[:iters] if trace else None

    def stage_times(self, X, reps=5, frame=0):
        """Average device ms per batched stage (HIP events on the current stream):
        pre_z, fft_y, fft_x, gather, scatter_zero, scatter, ifft_x, ifft_y, post_z."""
        names = ["pre_z", "fft_y", "fft_x", "gather", "scatter_zero", "scatter", "ifft_x", "ifft_y",
                 "post_z"]
        ms = (ctypes.c_float * len(names))()
        self._c(lib().mrn_stage_times(self._h, frame, _ptr(X), int(reps), ms, _stream()))
        return dict(zip(names, list(ms)))

    def power_method(self, iters=30):
        step = ctypes.c_double()
        self._c(lib().mrn_power_method(self._h, int(iters), ctypes.byref(step), _stream()))
        return step.value

    def fista(self, iters, step, X=None, graph=True):
        dev = torch.device("cuda")
        if X is None:
            X = torch.empty((self.F,) + self.image_shape, dtype=torch.complex64, device=dev)
        res = (ctypes.c_double * max(iters, 1))()
        self._c(lib().mrn_fista(self._h, _ptr(X), int(iters), float(step), int(bool(graph)), res, _stream()))
        return X, torch.tensor(list(res), dtype=torch.float64)[:iters]

    # ------------------------------------------------------ L1 wavelet step
    def wavelet(self, X, levels=(4, 4, 4), shift=(0, 0, 0), inverse=False):
        """In-place 3-D Daubechies-4 wavelet of X (F, Nz, Ny, Nx) or (Nz, Ny, Nx)."""
        frames = X.shape[0] if X.dim() == 4 else 1
        _need(X, torch.complex64, (frames,) + self.image_shape if X.dim() == 4 else self.image_shape, "X")
        lv = (ctypes.c_int32 * 3)(*levels)
        sh = (ctypes.c_int32 * 3)(*shift)
        self._c(lib().mrn_wavelet(self._h, _ptr(X), frames, lv, sh, int(bool(inverse)), _stream()))
        return X

    def l1_wavelet(self, X, fraction, levels=(4, 4, 4), spins=4, shifts=None):
        """The L1 wavelet proximal step on X (F, Nz, Ny, Nx), in place; returns
        the applied threshold.  shifts: (spins, 3) ints in 0..3 or None."""
        frames = X.shape[0] if X.dim() == 4 else 1
        lv = (ctypes.c_int32 * 3)(*levels)
        sh = None
        if shifts is not None:
            flat = [int(v) for row in shifts for v in row]
            sh = (ctypes.c_int32 * len(flat))(*flat)
            spins = len(flat) // 3
        thr = ctypes.c_double()
        self._c(lib().mrn_l1_wavelet(self._h, _ptr(X), frames, float(fraction), lv, int(spins), sh,
                                     ctypes.byref(thr), _stream()))
        return thr.value

    def set_fista_l1(self, fraction, levels=(4, 4, 4), spins=4, seed=0):
        """Apply the L1 wavelet step after each FISTA gradient step (0 = off)."""
        lv = (ctypes.c_int32 * 3)(*levels)
        self._c(lib().mrn_set_fista_l1(self._h, float(fraction), lv, int(spins), ctypes.c_uint64(seed)))

    def set_fista_llr(self, params=None, **kw):
        """Apply the locally low-rank threshold over the frames inside each FISTA
        iteration (before the L1 step); ``params`` from ``llr.llr_params`` or its
        keywords; alpha_time 0 (or params None with no keywords) turns it off."""
        from .llr import llr_params
        if params is None and kw:
            params = llr_params(**kw)
        self._c(lib().mrn_set_fista_llr(self._h, ctypes.byref(params) if params is not None else None))

    def fista_llr_smax(self):
        out = ctypes.c_float()
        self._c(lib().mrn_get_fista_llr_smax(self._h, ctypes.byref(out), _stream()))
        return out.value

    def set_fista_step(self, step_type="maxeig", cauchy_update_number=30):
        """FISTA step rule: "maxeig" (the step fista() is given) or "cauchy"
        (recomputed from the gradient for the first cauchy_update_number
        iterations, recon_lib.cpp:2013-2041)."""
        t = {"cauchy": 0, "maxeig": 1}[step_type]
        self._c(lib().mrn_set_fista_step(self._h, t, int(cauchy_update_number)))

    def fista_steps(self, n):
        """Complex step of each of the first n iterations of the last Cauchy FISTA."""
        out = (ctypes.c_double * max(2 * n, 2))()
        self._c(lib().mrn_get_fista_steps(self._h, out, int(n), _stream()))
        v = list(out)[: 2 * n]
        return torch.tensor([complex(v[2 * i], v[2 * i + 1]) for i in range(n)], dtype=torch.complex128)

    def fista_thresholds(self, n):
        out = (ctypes.c_double * max(n, 1))()
        self._c(lib().mrn_get_fista_thresholds(self._h, out, int(n), _stream()))
        return torch.tensor(list(out)[:n], dtype=torch.float64)

    # ------------------------------------------------------------- gradient
    def traj_grad(self, X, g, frame=0):
        """grad_k[m, a] = Re sum_c conj(g[c, m]) dA_c X[m]/dk_a  -> (M, 3) float32."""
        _need(X, torch.complex64, self.image_shape, "X")
        _need(g, torch.complex64, (self.C, self.M[frame]), "g")
        out = torch.empty((self.M[frame], 3), dtype=torch.float32, device=X.device)
        self._c(lib().mrn_traj_grad(self._h, frame, _ptr(X), _ptr(g), _ptr(out), _stream()))
        return out

    # ------------------------------------------------ grid-level stages
    def grid(self):
        """The op's oversampled grids as a torch view (no copy), shaped
        (C, Sz, Sy, Sx) for the caller: the reference's k3d_grid per coil.
        Internally they are stored coil-interleaved; this returns a strided
        view of that memory, so writes through it reach the op."""
        ptr, dims = ctypes.c_void_p(), (ctypes.c_int64 * 5)()
        self._c(lib().mrn_get_grid(self._h, ctypes.byref(ptr), dims))
        nch, CC, Sx, Sy, Sz = [int(v) for v in dims]
        n = nch * CC * Sx * Sy * Sz
        mem = _device_buffer(ptr.value, n)
        g = mem.view(nch, Sz, Sy, Sx, CC).permute(0, 4, 1, 2, 3).reshape(nch * CC, Sz, Sy, Sx)
        return g[: self.C] if nch * CC >= self.C else g

    def grid_raw(self):
        """(nch, Sz, Sy, Sx, CC) view of the grid memory."""
        ptr, dims = ctypes.c_void_p(), (ctypes.c_int64 * 5)()
        self._c(lib().mrn_get_grid(self._h, ctypes.byref(ptr), dims))
        nch, CC, Sx, Sy, Sz = [int(v) for v in dims]
        return _device_buffer(ptr.value, nch * CC * Sx * Sy * Sz).view(nch, Sz, Sy, Sx, CC)

    def grid_zero(self):
        self._c(lib().mrn_grid_zero(self._h, _stream()))

    def chop_grid_forward(self, data, frame=0):
        """grid_c += scatter(kw taper data_c) (gridFFT::chop_grid_forward)."""
        _need(data, torch.complex64, (self.C, self.M[frame]), "data")
        self._c(lib().mrn_chop_grid_forward(self._h, frame, _ptr(data), _stream()))

    def chop_grid_backward(self, data=None, frame=0, sub=None):
        """data_c = gather(grid_c) (gridFFT::chop_grid_backward); returns data."""
        if data is None:
            data = torch.empty((self.C, self.M[frame]), dtype=torch.complex64, device="cuda")
        _need(data, torch.complex64, (self.C, self.M[frame]), "data")
        if sub is not None:
            _need(sub, torch.complex64, (self.C, self.M[frame]), "sub")
        self._c(lib().mrn_chop_grid_backward(self._h, frame, _ptr(data), _ptr(sub), _stream()))
        return data

    def sync(self):
        self._c(lib().mrn_sync(self._h, _stream()))
