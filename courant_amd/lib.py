"""ctypes binding of the courant.h C ABI and a Sim class shaped like the
TypeScript wrapper's (web/src/solver/courant.ts:121-455).

The same binding drives any library that exports courant.h: this engine's
libcourant.so, or (in tests only) a CPU checker library with the same ABI.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIB = os.path.join(HERE, "libcourant.so")

C0 = 299792458.0
MU0 = 1.25663706212e-6
EPS0 = 8.8541878128e-12
X, Y, Z = 0, 1, 2
XLO, XHI, YLO, YHI, ZLO, ZHI = range(6)
PEC, CPML = 0, 1

_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_L = C.c_long
_S = C.c_size_t
_DP = C.POINTER(C.c_double)
_UP = C.POINTER(C.c_ubyte)
_IP = C.POINTER(C.c_int)

# name: (restype, argtypes) for every courant.h entry point
SIGNATURES = {
    "cem_create": (_P, [_I, _I, _I, _DP, _DP, _DP]),
    "cem_destroy": (None, [_P]),
    "cem_last_error": (C.c_char_p, []),
    "cem_version": (C.c_char_p, []),
    "cem_set_boundary": (_I, [_P, _I, _I, _I]),
    "cem_material_box": (_I, [_P, _D, _D, _D, _D, _D, _D, _D, _D]),
    "cem_set_debye_poles": (_I, [_P, _I, _DP]),
    "cem_material_box_debye": (_I, [_P, _D, _D, _D, _D, _D, _D, _D, _D, _DP]),
    "cem_set_conductor_loss": (_I, [_P, _D, _I, _D, _D, _D, _D]),
    "cem_conductor_fit_error": (_D, [_P]),
    "cem_conductor_zs": (None, [_P, _D, _DP, _DP]),
    "cem_fit_djordjevic_sarkar": (_D, [_D, _D, _D, _D, _D, _I, _DP, _DP, _DP]),
    "cem_pec_box": (_I, [_P, _D, _D, _D, _D, _D, _D]),
    "cem_pec_sheet": (_I, [_P, _I, _D, _D, _D, _D, _D]),
    "cem_pec_edges_plane": (_I, [_P, _I, _I, _UP, _UP]),
    "cem_pec_cells": (_I, [_P, _I, _I, _I, _I, _I, _I, _UP]),
    "cem_add_lumped_port": (_I, [_P, _I, _I, _I, _I, _I, _I, _D]),
    "cem_set_gaussian": (_I, [_P, _I, _D, _D, _D]),
    "cem_set_frequencies": (_I, [_P, _I, _DP]),
    "cem_add_vprobe": (_I, [_P, _I, _I, _I, _I, _I, _I]),
    "cem_add_iprobe": (_I, [_P, _I, _I, _I, _I, _I, _I]),
    "cem_finalize": (_I, [_P, _D]),
    "cem_dt": (_D, [_P]),
    "cem_step_count": (_L, [_P]),
    "cem_step": (None, [_P, _L]),
    "cem_run": (_L, [_P, _L, _D, _I]),
    "cem_energy": (_D, [_P]),
    "cem_nfreq": (_I, [_P]),
    "cem_port_v": (_DP, [_P, _I]),
    "cem_port_i": (_DP, [_P, _I]),
    "cem_source_spectrum": (_DP, [_P]),
    "cem_vprobe": (_DP, [_P, _I]),
    "cem_iprobe": (_DP, [_P, _I]),
    "cem_port_impedance": (_D, [_P, _I]),
    "cem_excited_port": (_I, [_P]),
    "cem_field": (_DP, [_P, _I]),
    "cem_index": (_S, [_P, _I, _I, _I]),
    "cem_dims": (None, [_P, _IP, _IP, _IP]),
    "cem_mode2d_create": (_P, [_I, _I, _DP, _DP]),
    "cem_mode2d_destroy": (None, [_P]),
    "cem_mode2d_set_media": (None, [_P, _DP, _DP, _DP, _UP, _UP, _UP]),
    "cem_mode2d_solve": (_I, [_P, _D, _D, _DP, _I, _D]),
    "cem_mode2d_set_debye": (None, [_P, _I, _DP, _DP, _DP, _DP]),
    "cem_mode2d_set_loss": (None, [_P, _DP, _DP, _DP]),
    "cem_mode2d_set_pml": (None, [_P, _I, _DP, _DP, _DP]),
    "cem_mode2d_beta": (_D, [_P]),
    "cem_mode2d_alpha": (_D, [_P]),
    "cem_mode2d_eu": (_DP, [_P]),
    "cem_mode2d_ev": (_DP, [_P]),
    "cem_mode2d_hu": (_DP, [_P]),
    "cem_mode2d_hv": (_DP, [_P]),
    "cem_mode2d_power": (_D, [_P]),
    "cem_mode2d_conductor_alpha": (_D, [_P, _D]),
    "cem_mode2d_vi": (None, [_P, _I, _I, _I, _I, _I, _I, _I, _DP, _DP]),
    "cem_add_mode_port": (_I, [_P, _P]),
    "cem_add_mode_port_args": (_I, [_P] + [_I] * 15 + [_D, _D, _D]),
    "cem_set_mode_excitation": (_I, [_P, _I, _D, _D]),
    "cem_mode_port_waves": (_I, [_P, _I, _DP, _DP, _DP, _DP]),
    "cem_mode_port_alpha": (_I, [_P, _I, _DP]),
    "cem_mode_port_alpha_c": (_I, [_P, _I, _DP]),
    "cem_s2_renormalize": (None, [_DP, _D, _D, _D]),
    "cem_plan_array": (_DP, [_P, _I, _I]),
    "cem_plan_npml": (_I, [_P, _I]),
    "cem_plan_has_pml": (_I, [_P]),
    "cem_plan_nport": (_I, [_P]),
    "cem_plan_nvprobe": (_I, [_P]),
    "cem_plan_niprobe": (_I, [_P]),
    "cem_plan_ndisp": (_I, [_P]),
    "cem_plan_nsibc": (_I, [_P]),
    "cem_plan_port": (_I, [_P, _I, _DP, _DP, _DP]),
    "cem_plan_vprobe": (_I, [_P, _I, _DP, _DP]),
    "cem_plan_iprobe": (_I, [_P, _I, _DP, _DP]),
    "cem_plan_mode_source": (_I, [_P, _DP, _DP]),
    "cem_plan_waveforms": (None, [_P, _L, _I, _DP, _DP]),
    "cem_plan_excited_lumped": (_I, [_P]),
    "cem_plan_npole": (_I, [_P]),
    "cem_plan_disp": (_I, [_P, _DP, _DP, _DP, _DP]),
    "cem_plan_sibc": (_I, [_P, _DP, _DP, _DP, _DP, _DP, _IP]),
    "cem_push_samples": (None, [_P, _I, _DP, _DP, _DP, _DP]),
    "cem_checkpoint_size": (_S, [_P]),
    "cem_checkpoint_save": (_I, [_P, _DP]),
    "cem_checkpoint_load": (_I, [_P, _DP]),
    "cem_state_array": (_DP, [_P, _I, _I, _I]),
    "cem_state_len": (_S, [_P, _I, _I, _I]),
    "cem_mt_pre": (None, [_P]),
    "cem_mt_H": (None, [_P, _I, _I]),
    "cem_mt_mid": (None, [_P]),
    "cem_mt_E": (None, [_P, _I, _I]),
    "cem_mt_post": (None, [_P]),
    "cem_set_step": (None, [_P, _L]),
    "cem_cross_section_x": (_P, [_P, _I, _I, _I, _I, _I]),
}

# engine extensions (include/courant_amd.h)
AMD_SIGNATURES = {
    "cem_amd_kernel_stats": (_I, [_P, _I, _DP, C.POINTER(C.c_long)]),
    "cem_amd_set_profiling": (_I, [_P, _I]),
    "cem_amd_synchronize": (_I, [_P]),
    "cem_amd_step_device": (_I, [_P, _L]),
    "cem_amd_tune": (_L, [_P]),
    "cem_amd_flush_dft": (_I, [_P]),
    "cem_amd_copy_bandwidth": (_D, [_S, _I]),
    "cem_amd_s2_renormalize": (_I, [_DP, _I, _DP, _DP, _D]),
    "cem_amd_s2_renormalize_noalloc": (_I, [_DP, _I, _DP, _DP, _D]),
    "cem_amd_launch_shape": (_I, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cem_amd_lane_widths": (_I, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cem_amd_tile_order": (_I, [_P]),
    "cem_amd_class_bytes": (_I, [_P]),
    "cem_amd_step_shape": (_I, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cem_amd_clone_excitation": (_P, [_P, _I]),
    "cem_amd_step_pair": (_I, [_P, _P, _L]),
}


class ModePortDef(C.Structure):
    """cem_mode_port_def: 15 ints then 3 doubles (courant.h:198-208)."""
    _fields_ = [(n, C.c_int) for n in ("dir", "i_src", "i_mon", "j0", "j1", "k0", "k1", "vj", "vk0", "vk1",
                                        "vsign", "ja", "jb", "ka", "kb")] + \
               [("x_ref", C.c_double), ("f0", C.c_double), ("neff0", C.c_double)]


class CourantError(RuntimeError):
    pass


def _dp(a):
    return a.ctypes.data_as(_DP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class _View(np.ndarray):
    """A numpy view of library-owned memory that keeps the owning Sim alive (the
    memory is freed when the Sim is closed or collected)."""
    pass


def _owned_view(ptr, n, owner):
    v = np.ctypeslib.as_array(ptr, shape=(n,)).view(_View)
    v._owner = owner
    return v


class Courant:
    """A loaded courant.h library (product engine or a checker with the same ABI)."""

    def __init__(self, path=DEFAULT_LIB, extensions=None):
        if not os.path.exists(path):
            raise CourantError(f"courant library not built: {path} (run __graft_entry__.build())")
        self.path = path
        self.lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(self.lib, name)
            fn.restype = res
            fn.argtypes = args
        self.has_amd = extensions if extensions is not None else hasattr(self.lib, "cem_amd_step_device")
        if self.has_amd:
            for name, (res, args) in AMD_SIGNATURES.items():
                fn = getattr(self.lib, name)
                fn.restype = res
                fn.argtypes = args

    def last_error(self):
        return self.lib.cem_last_error().decode()

    def version(self):
        return self.lib.cem_version().decode()

    def create_sim(self, x, y, z):
        return Sim(self, x, y, z)

    def fit_djordjevic_sarkar(self, dk, df, fref, fmin, fmax, npole):
        einf = C.c_double()
        fp = np.zeros(npole)
        deps = np.zeros(npole)
        err = self.lib.cem_fit_djordjevic_sarkar(dk, df, fref, fmin, fmax, npole, C.byref(einf), _dp(fp), _dp(deps))
        return {"eps_inf": einf.value, "fp": fp, "deps": deps, "tan_delta_error": err}

    def renormalize_batch(self, S, z1, z2, zref):
        """S: nf x 2 x 2 complex, z1, z2: nf reference impedances; renormalised on the GPU."""
        S = np.asarray(S, complex)
        nf = S.shape[0]
        v = np.ascontiguousarray(np.stack([S.real, S.imag], -1).reshape(nf, 8))
        rc = self.lib.cem_amd_s2_renormalize(_dp(v), nf, _dp(_f64(z1)), _dp(_f64(z2)), zref)
        if rc < 0:
            raise CourantError(self.last_error())
        v = v.reshape(nf, 2, 2, 2)
        return v[..., 0] + 1j * v[..., 1]

    def renormalize2(self, S, z1, z2, zref):
        """S: 4 complex values (S11, S12, S21, S22); returns the renormalised 4."""
        v = np.zeros(8)
        for q, s in enumerate(S):
            v[2 * q], v[2 * q + 1] = s.real, s.imag
        self.lib.cem_s2_renormalize(_dp(v), z1, z2, zref)
        return [complex(v[2 * q], v[2 * q + 1]) for q in range(4)]


class Sim:
    """One simulation handle; method names follow courant.ts's Sim."""

    def __init__(self, cem, x, y, z, handle=None):
        self.cem = cem
        self.lib = cem.lib
        self.x, self.y, self.z = _f64(x), _f64(y), _f64(z)
        self.n = (len(self.x) - 1, len(self.y) - 1, len(self.z) - 1)
        h = handle or self.lib.cem_create(self.n[0], self.n[1], self.n[2], _dp(self.x), _dp(self.y), _dp(self.z))
        if not h:
            raise CourantError(cem.last_error())
        self.h = h
        self.N = (self.n[0] + 1) * (self.n[1] + 1) * (self.n[2] + 1)
        self.freqs = np.array([1e9])          # cem_finalize's default (sim.c:429)
        self.excitation_lumped = (0.0, 0.0, 1.0)
        self.excitation_mode = None

    def close(self):
        if self.h:
            self.lib.cem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ok(self, rc):
        if rc < 0:
            raise CourantError(self.cem.last_error())
        return rc

    # ---- set-up ----
    def set_boundary(self, face, kind, n=0):
        return self._ok(self.lib.cem_set_boundary(self.h, face, kind, n))

    def material_box(self, x0, x1, y0, y1, z0, z1, er, sigma=0.0):
        return self._ok(self.lib.cem_material_box(self.h, x0, x1, y0, y1, z0, z1, er, sigma))

    def set_debye_poles(self, fp):
        fp = _f64(fp)
        return self._ok(self.lib.cem_set_debye_poles(self.h, len(fp), _dp(fp)))

    def material_box_debye(self, x0, x1, y0, y1, z0, z1, eps_inf, sigma, deps):
        deps = _f64(deps)
        return self._ok(self.lib.cem_material_box_debye(self.h, x0, x1, y0, y1, z0, z1, eps_inf, sigma, _dp(deps)))

    def set_conductor_loss(self, sigma, rough, p1, p2, fmin, fmax):
        return self._ok(self.lib.cem_set_conductor_loss(self.h, sigma, rough, p1, p2, fmin, fmax))

    def conductor_fit_error(self):
        return self.lib.cem_conductor_fit_error(self.h)

    def pec_box(self, x0, x1, y0, y1, z0, z1):
        return self._ok(self.lib.cem_pec_box(self.h, x0, x1, y0, y1, z0, z1))

    def pec_sheet(self, axis, pos, a0, a1, b0, b1):
        return self._ok(self.lib.cem_pec_sheet(self.h, axis, pos, a0, a1, b0, b1))

    def pec_edges_plane(self, axis, m, mask_u=None, mask_v=None):
        mu = None if mask_u is None else np.ascontiguousarray(mask_u, dtype=np.uint8)
        mv = None if mask_v is None else np.ascontiguousarray(mask_v, dtype=np.uint8)
        return self._ok(self.lib.cem_pec_edges_plane(self.h, axis, m,
                                                     None if mu is None else mu.ctypes.data_as(_UP),
                                                     None if mv is None else mv.ctypes.data_as(_UP)))

    def pec_cells(self, i0, j0, k0, mask):
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        ni, nj, nk = mask.shape
        return self._ok(self.lib.cem_pec_cells(self.h, i0, j0, k0, ni, nj, nk, mask.ctypes.data_as(_UP)))

    def add_lumped_port(self, axis, n1, n2, m0, m1, direction, R):
        return self._ok(self.lib.cem_add_lumped_port(self.h, axis, n1, n2, m0, m1, direction, R))

    def set_gaussian(self, port, amp, t0, tau):
        rc = self._ok(self.lib.cem_set_gaussian(self.h, port, amp, t0, tau))
        self.excitation_lumped = (amp, t0, tau)
        return rc

    def set_frequencies(self, f):
        f = _f64(f)
        self.freqs = f
        return self._ok(self.lib.cem_set_frequencies(self.h, len(f), _dp(f)))

    def add_vprobe(self, axis, n1, n2, m0, m1, direction):
        return self._ok(self.lib.cem_add_vprobe(self.h, axis, n1, n2, m0, m1, direction))

    def add_iprobe(self, axis, m, a0, a1, b0, b1):
        return self._ok(self.lib.cem_add_iprobe(self.h, axis, m, a0, a1, b0, b1))

    def add_mode_port(self, **kw):
        d = ModePortDef(**kw)
        return self._ok(self.lib.cem_add_mode_port(self.h, C.cast(C.pointer(d), _P)))

    def set_mode_excitation(self, port, t0, tau):
        rc = self._ok(self.lib.cem_set_mode_excitation(self.h, port, t0, tau))
        self.excitation_mode = (t0, tau)
        return rc

    def finalize(self, cfl=0.99):
        return self._ok(self.lib.cem_finalize(self.h, cfl))

    # ---- running ----
    @property
    def dt(self):
        return self.lib.cem_dt(self.h)

    @property
    def steps(self):
        return self.lib.cem_step_count(self.h)

    def step(self, n):
        self.lib.cem_step(self.h, n)

    def run(self, max_steps, decay, check_every=500):
        return self.lib.cem_run(self.h, max_steps, decay, check_every)

    def energy(self):
        return self.lib.cem_energy(self.h)

    # ---- results ----
    @property
    def nf(self):
        return self.lib.cem_nfreq(self.h)

    def _spectrum(self, ptr):
        if not ptr:
            raise CourantError(self.cem.last_error() or "no spectrum")
        a = np.ctypeslib.as_array(ptr, shape=(2 * self.nf,))
        return a[0::2] + 1j * a[1::2]

    def port_v(self, p):
        return self._spectrum(self.lib.cem_port_v(self.h, p))

    def port_i(self, p):
        return self._spectrum(self.lib.cem_port_i(self.h, p))

    def source_spectrum(self):
        return self._spectrum(self.lib.cem_source_spectrum(self.h))

    def vprobe(self, p):
        return self._spectrum(self.lib.cem_vprobe(self.h, p))

    def iprobe(self, p):
        return self._spectrum(self.lib.cem_iprobe(self.h, p))

    def field(self, comp):
        """Live FP64 view of field component comp (0..2 E, 3..5 H), shape (nx+1, ny+1, nz+1)."""
        ptr = self.lib.cem_field(self.h, comp)
        if not ptr:
            raise CourantError(self.cem.last_error() or "cem_field returned NULL")
        return _owned_view(ptr, self.N, self).reshape(self.n[0] + 1, self.n[1] + 1, self.n[2] + 1)

    def index(self, i, j, k):
        return self.lib.cem_index(self.h, i, j, k)

    def mode_port_waves(self, port):
        nf = self.nf
        a, b, z0, beta = np.zeros(2 * nf), np.zeros(2 * nf), np.zeros(nf), np.zeros(nf)
        self._ok(self.lib.cem_mode_port_waves(self.h, port, _dp(a), _dp(b), _dp(z0), _dp(beta)))
        return {"a": a[0::2] + 1j * a[1::2], "b": b[0::2] + 1j * b[1::2], "z0": z0, "beta": beta}

    def mode_port_alpha(self, port, conductor=False):
        out = np.zeros(self.nf)
        fn = self.lib.cem_mode_port_alpha_c if conductor else self.lib.cem_mode_port_alpha
        self._ok(fn(self.h, port, _dp(out)))
        return out

    # ---- plan (export.c) ----
    def plan(self):
        """Finalized update plan as numpy arrays (mirrors courant.ts Sim.plan)."""
        L = self.lib
        h = self.h
        nx, ny, nz = self.n
        lens = {0: self.N, 1: self.N, 2: None, 3: None, 4: None, 5: None, 6: None, 7: None, 8: None, 9: None}
        out = {"dt": self.dt, "N": self.N, "n": self.n}
        names = ["ca", "cb", "ide", "idh", "be", "ce", "bh", "ch", "d", "dd"]
        for w, nm in enumerate(names):
            arrs = []
            for c in range(3):
                nn = self.n[c]
                ln = lens[w] or (nn + 1 if nm in ("ide", "be", "ce", "dd") else nn)
                ptr = L.cem_plan_array(h, w, c)
                arrs.append(np.ctypeslib.as_array(ptr, shape=(ln,)).copy())
            out[nm] = arrs
        out["npml"] = [L.cem_plan_npml(h, f) for f in range(6)]
        out["has_pml"] = L.cem_plan_has_pml(h)
        ports = []
        for p in range(L.cem_plan_nport(h)):
            n = L.cem_plan_port(h, p, None, None, None)
            idx, vw, cs = np.zeros(n), np.zeros(n), np.zeros(n)
            L.cem_plan_port(h, p, _dp(idx), _dp(vw), _dp(cs))
            ports.append({"idx": idx, "vw": vw, "cs": cs})
        out["ports"] = ports
        for kind, fn, cnt in (("vprobes", L.cem_plan_vprobe, L.cem_plan_nvprobe),
                              ("iprobes", L.cem_plan_iprobe, L.cem_plan_niprobe)):
            lst = []
            for p in range(cnt(h)):
                n = fn(h, p, None, None)
                idx, w = np.zeros(n), np.zeros(n)
                fn(h, p, _dp(idx), _dp(w))
                lst.append({"idx": idx, "w": w})
            out[kind] = lst
        n = L.cem_plan_mode_source(h, None, None)
        idx, w = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        L.cem_plan_mode_source(h, _dp(idx), _dp(w))
        out["mode_source"] = {"idx": idx[:n], "w": w[:n]}
        out["excited_lumped"] = L.cem_plan_excited_lumped(h)
        nd, npole = L.cem_plan_ndisp(h), L.cem_plan_npole(h)
        if nd:
            idx, cb, beta, k = np.zeros(nd), np.zeros(nd), np.zeros(nd * npole), np.zeros(npole)
            L.cem_plan_disp(h, _dp(idx), _dp(cb), _dp(beta), _dp(k))
            out["disp"] = {"npole": npole, "idx": idx, "cb": cb, "beta": beta, "k": k}
        ns = L.cem_plan_nsibc(h)
        if ns:
            idx, hidx, hw = np.zeros(ns), np.zeros(4 * ns), np.zeros(4 * ns)
            c, e = np.zeros(16), np.zeros(16)
            nt = C.c_int()
            L.cem_plan_sibc(h, _dp(idx), _dp(hidx), _dp(hw), _dp(c), _dp(e), C.byref(nt))
            out["sibc"] = {"idx": idx, "hidx": hidx, "hw": hw, "c": c[:nt.value], "e": e[:nt.value]}
        out["ndisp"], out["nsibc"] = nd, ns
        out["state_len"] = {(w, c, k): L.cem_state_len(h, w, c, k) for w in range(5) for c in range(3)
                            for k in range(2)}
        return out

    def waveforms(self, n0, n):
        lumped, mode = np.zeros(n), np.zeros(n)
        self.lib.cem_plan_waveforms(self.h, n0, n, _dp(lumped), _dp(mode))
        return lumped, mode

    # ---- checkpoints ----
    def checkpoint(self):
        n = self.lib.cem_checkpoint_size(self.h)
        buf = np.zeros(n)
        self._ok(self.lib.cem_checkpoint_save(self.h, _dp(buf)))
        return buf

    def restore(self, buf):
        buf = _f64(buf)
        return self._ok(self.lib.cem_checkpoint_load(self.h, _dp(buf)))

    # ---- engine extensions ----
    def step_device(self, n):
        """Advance n steps on the GPU without refreshing host mirrors (timed loops)."""
        return self._ok(self.lib.cem_amd_step_device(self.h, n))

    def synchronize(self):
        return self._ok(self.lib.cem_amd_synchronize(self.h))

    def clone_excitation(self, port):
        """A simulation with this one's plan for another excited port (lumped port `port`
        with this one's Gaussian, or mode port `port` when this one drives a mode port): no
        set-up step runs again; its fields start from zero at step 0."""
        h = self.lib.cem_amd_clone_excitation(self.h, port)
        if not h:
            raise CourantError(self.cem.last_error())
        c = Sim(self.cem, self.x, self.y, self.z, handle=h)
        c.freqs = self.freqs.copy()
        c.excitation_lumped = self.excitation_lumped
        c.excitation_mode = self.excitation_mode
        return c

    def step_pair(self, other, n):
        """Advance this simulation and its excitation clone `other` (same step) n steps: the H
        and E updates of both run as one launch each per step."""
        return self._ok(self.lib.cem_amd_step_pair(self.h, other.h, n))

    def tune(self):
        """Run the launch-shape autotuner now; returns the steps it ran."""
        return self._ok(self.lib.cem_amd_tune(self.h))

    def flush_dft(self):
        """Queue the DFT of the samples buffered so far (no host transfer)."""
        return self._ok(self.lib.cem_amd_flush_dft(self.h))

    def set_profiling(self, on):
        return self._ok(self.lib.cem_amd_set_profiling(self.h, 1 if on else 0))

    def launch_shape(self):
        """(i-planes per thread of the H update, of the E update, fixed yet)."""
        h, e = C.c_int(), C.c_int()
        rc = self._ok(self.lib.cem_amd_launch_shape(self.h, C.byref(h), C.byref(e)))
        return h.value, e.value, rc == 1

    def lane_widths(self):
        """(k points per thread of the H update, of the E update, fixed yet)."""
        h, e = C.c_int(), C.c_int()
        rc = self._ok(self.lib.cem_amd_lane_widths(self.h, C.byref(h), C.byref(e)))
        return h.value, e.value, rc == 1

    def tile_order(self):
        """"xcd" (one band of tiles per XCD), "plain" (dispatch order) or "plain<T>" (dispatch
        order in chunks of T tiles); the value COURANT_TILE_ORDER takes."""
        v = self._ok(self.lib.cem_amd_tile_order(self.h))
        if v & 1:
            return "xcd"
        return "plain" if v < 4 else f"plain{1 << (v >> 2)}"

    def h_reversed(self):
        """True when the H update runs its tiles from the last (zig-zag with the E update)."""
        return bool(self._ok(self.lib.cem_amd_tile_order(self.h)) & 2)

    def class_bytes(self):
        """Bytes per grid point of the coefficient data the E update reads (1, 4 or 24)."""
        return self._ok(self.lib.cem_amd_class_bytes(self.h))

    def step_shape(self):
        """(grid rows per workgroup, x planes per workgroup) of the one-pass step kernel;
        (0, 0) when full steps run the separate H and E updates."""
        r, p = C.c_int(), C.c_int()
        self._ok(self.lib.cem_amd_step_shape(self.h, C.byref(r), C.byref(p)))
        return r.value, p.value

    def kernel_stats(self, which):
        """(total GPU ms, launches) of kernel `which`: 0 H update, 1 E update, 2 gather, 3 DFT, 4 post,
        5 one-pass step (H and E in one kernel)."""
        t, n = C.c_double(), C.c_long()
        self._ok(self.lib.cem_amd_kernel_stats(self.h, which, C.byref(t), C.byref(n)))
        return t.value, n.value


def s_params(v1, i1, v2, i2, z=50.0):
    """Port-1-excited S11, S21 from lumped V/I spectra (gcpw_line.c:379-391, courant.ts:458-470)."""
    sq = np.sqrt(z)
    a1 = (v1 + z * i1) / (2 * sq)
    b1 = (v1 - z * i1) / (2 * sq)
    b2 = (v2 - z * i2) / (2 * sq)
    return b1 / a1, b2 / a1
