"""cubesphere_amd -- MI355X (gfx950) cell-update path of the CubeSphere solver.

Host-side mirror of the reference's per-patch driver interface
(_I/_1000riemann_geomesh_cs_mhd/1000riemann_geomesh_cs_mhd.f:2007-3120) over the
C-ABI in include/cubesphere_hip.h.  The library is loaded with ctypes from this
package directory; there is no fallback: if libcubesphere_hip.so is missing or
no GPU is present the device entry points raise.
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libcubesphere_hip.so")

RECON_PCM, RECON_WENO5, RECON_P1ADP = 0, 1, 2
RS_HLL, RS_HLLD = 0, 1
EMF_LLF, EMF_HLL = 0, 1
FP64, MIXED = 0, 1
RBC_PAD, RBC_WALL, RBC_PROBLEM = 0, 1, 2
RMESH_UNIFORM, RMESH_LOG = 0, 1
# the launch groups of the P1 stage (CS_P1K_* of the header, in order)
P1_KERNEL_NAMES = ["exchange", "prim", "face_sync", "flatten", "weno", "positivity", "replay", "face_points",
                   "clsq", "flux", "edge", "sync", "update_faces", "update_zones", "imex", "cfl"]
KERNEL_NAMES = ["halo", "prims", "flux_x", "flux_y", "flux_z", "emf_z", "emf_x", "emf_y",
                "update_cells", "update_faces", "cleanup", "cfl", "step", "sync", "flux_xy", "emf"]
(K_HALO, K_PRIMS, K_FLUX_X, K_FLUX_Y, K_FLUX_Z, K_EMF_Z, K_EMF_X, K_EMF_Y, K_UPDATE_CELLS,
 K_UPDATE_FACES, K_CLEANUP, K_CFL, K_STEP, K_SYNC, K_FLUX_XY, K_EMF) = range(16)


class Config(ctypes.Structure):
    """Mirror of cs_config."""
    _fields_ = ([(n, ctypes.c_int32) for n in
                 ("ntp", "nr", "nghost", "recon", "riemann", "emf", "precision",
                  "rk_order", "radial_bc", "radial_mesh")] +
                [(n, ctypes.c_double) for n in
                 ("rmin", "rmax", "gamma", "rhofloor", "prsfloor", "cfl", "grav")] +
                [("imex", ctypes.c_int32), ("bc_problem", ctypes.c_int32),
                 ("omega", ctypes.c_double * 3)])

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Sizes(ctypes.Structure):
    _fields_ = [("n_cell", ctypes.c_int64), ("n_face", ctypes.c_int64),
                ("n_vert2d", ctypes.c_int64), ("n_interior", ctypes.c_int64)]


def make_config(ntp, nr=1, nghost=3, recon=RECON_WENO5, riemann=RS_HLLD, emf=EMF_LLF,
                precision=FP64, rk_order=3, radial_bc=None, radial_mesh=RMESH_UNIFORM,
                rmin=1.0, rmax=None, gamma=5.0 / 3.0, rhofloor=1e-12, prsfloor=1e-9,
                cfl=0.4, grav=0.0, imex=0, omega=(0.0, 0.0, 0.0), bc_problem=0):
    """Build a Config.  nr == 1 is the thin-shell surface mode: rmax defaults
    to rmin + (pi/2)/ntp and the radial boundaries are conducting walls.
    bc_problem: the problem whose state RBC_PROBLEM writes into the radial
    ghost layers (ADD_BOUNDARY)."""
    if rmax is None:
        rmax = rmin + (0.5 * math.pi) / ntp if nr == 1 else 1.5 * rmin
    if radial_bc is None:
        radial_bc = RBC_WALL if nr == 1 else RBC_PAD
    return Config(ntp, nr, nghost, recon, riemann, emf, precision, rk_order, radial_bc,
                  radial_mesh, rmin, rmax, gamma, rhofloor, prsfloor, cfl, grav, imex, bc_problem,
                  (ctypes.c_double * 3)(*omega))


# The five workloads of the benchmark description (sizes, schemes, ICs).
def workload(name, precision=FP64):
    """Return (config, problem id, steps) of a named workload."""
    if name == "C1":   # 6x32x32 surface, WENO5 + HLL + SSP-RK3, Gaussian bump + rotation
        return make_config(32, 1, 3, RECON_WENO5, RS_HLL, EMF_LLF, precision, 3), 1, 100
    if name == "C2":   # 6x256^2 surface, WENO-ADP (the reference's P1 scheme) + HLLD + SSP-RK3, CT, magnetised blast
        return make_config(256, 1, 4, RECON_P1ADP, RS_HLLD, EMF_LLF, precision, 3), 2, 500
    if name == "C2W5":  # the same surface with the stage path's own WENO5
        return make_config(256, 1, 3, RECON_WENO5, RS_HLLD, EMF_LLF, precision, 3), 2, 500
    if name == "C3":   # 6x192^2x64 shell r in [1,1.5], WENO5 + HLLD + CT, IMEX-RK with gravity 1/r^2 implicit
        return make_config(192, 64, 3, RECON_WENO5, RS_HLLD, EMF_LLF, precision, 2,
                           RBC_PAD, RMESH_UNIFORM, 1.0, 1.5, grav=1.0, imex=1), 3, 10
    if name == "C5":   # 6x512^2x128, WENO-ADP (the reference's P1 scheme) + HLLD + CT + IMEX-RK, gravity implicit
        return make_config(512, 128, 4, RECON_P1ADP, RS_HLLD, EMF_LLF, precision, 2,
                           RBC_PAD, RMESH_UNIFORM, 1.0, 1.5, grav=1.0, imex=1), 3, 50
    if name == "C5W5":  # the same shell with the stage path's own WENO5
        return make_config(512, 128, 3, RECON_WENO5, RS_HLLD, EMF_LLF, precision, 2,
                           RBC_PAD, RMESH_UNIFORM, 1.0, 1.5, grav=1.0, imex=1), 3, 50
    raise ValueError(name)


_lib = None


# status codes of the C-ABI (CS_E_*)
E_INVALID, E_BAD_DECOMP, E_HIP, E_NONPOSITIVE, E_NAN, E_NOMEM, E_IO, E_FORMAT = -1, -2, -3, -4, -5, -6, -7, -8


class CsError(RuntimeError):
    """A failed C-ABI call; code is its CS_E_* status."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def lib():
    """Load libcubesphere_hip.so (built by cubesphere_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise CsError(f"{LIBPATH} is missing: run `python -m cubesphere_amd.build`")
        # One HIP runtime per process: torch ships its own libamdhip64.so.7,
        # and a library loaded before it would bring in a second runtime that
        # leaves torch without devices.  Loading torch first makes the library
        # bind to the runtime already in the process.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIBPATH)
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        cfgp = ctypes.POINTER(Config)
        sig = {
            "cs_sizes_of": [cfgp, ctypes.POINTER(Sizes)],
            "cs_build_unit_lattice": [cfgp, vp],
            "cs_build_radial_mesh": [cfgp, vp],
            "cs_build_metric": [cfgp, vp, vp, vp, vp],
            "cs_init_problem": [cfgp, i32, ctypes.c_uint64, vp, vp, vp, vp, vp],
            "cs_create": [cfgp, ctypes.POINTER(vp)],
            "cs_destroy": [vp],
            "cs_upload_state": [vp, vp, vp, vp, vp, vp],
            "cs_download_state": [vp, vp, vp, vp, vp, vp],
            "cs_halo_exchange": [vp],
            "cs_halo_exchange_p1": [vp, vp, vp, vp, vp, vp, vp],
            "cs_make_facial_fluxes": [vp],
            "cs_make_edge_efields": [vp],
            "cs_update_stage": [vp, i32],
            "cs_evaluate_timestep": [vp],
            "cs_set_dt": [vp, dbl],
            "cs_get_dt": [vp, ctypes.POINTER(dbl)],
            "cs_rk_stage": [vp, i32],
            "cs_advance": [vp, i32, i32],
            "cs_download_fluxes": [vp, i32, vp, vp],
            "cs_download_emf": [vp, vp, vp, vp],
            "cs_conservation_sums": [vp, vp],
            "cs_max_divb": [vp, ctypes.POINTER(dbl)],
            "cs_time_steps": [vp, i32, i32, ctypes.POINTER(ctypes.c_float)],
            "cs_time_kernel": [vp, i32, i32, ctypes.POINTER(ctypes.c_float)],
            "cs_profile_steps": [vp, i32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(i32)],
            "cs_profile_steps_p1": [vp, i32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(i32)],
            "cs_init_problem_p1": [vp, i32, ctypes.c_uint64, ctypes.c_double],
            "cs_add_boundary_p1": [vp, i32, i32, i32, ctypes.c_uint64, ctypes.c_double],
            "cs_autotune": [vp, i32, ctypes.POINTER(i32)],
            "cs_geometry_sizes": [cfgp, vp],
            "cs_build_geometry": [cfgp, vp, vp],
            "cs_mode_indices": [i32, ctypes.POINTER(i32), vp, vp, vp],
            "cs_stencil_indexing": [cfgp, vp],
            "cs_limit_all_slopes_p1": [vp, vp],
            "cs_build_volume_p12": [cfgp, vp],
            "cs_limit_all_slopes_p12": [vp, vp],
            "cs_limit_all_slopes_p12_timed": [vp, vp, ctypes.POINTER(ctypes.c_float)],
            "cs_limit_all_slopes_p1_timed": [vp, vp, ctypes.POINTER(ctypes.c_float)],
            "cs_limit_all_slopes_p1_noalloc": [vp, vp],
            "cs_make_facial_fluxes_p1": [vp, vp],
            "cs_update_imex_rk_step_p1": [vp, vp],
            "cs_cleanup_after_update_p1": [vp, vp],
            "cs_make_edge_efields_p1": [vp, vp],
            "cs_build_non_stiff_sources_p1": [vp, vp],
            "cs_pad_boundary_r_p1": [vp, vp],
            "cs_sync_shared_p1": [vp, vp],
            "cs_evaluate_timestep_p1": [vp, ctypes.c_double, ctypes.c_double, vp, vp],
            "cs_batch_create": [vp, i32, ctypes.POINTER(vp)],
            "cs_batch_destroy": [vp],
            "cs_batch_shell": [vp, i32, ctypes.POINTER(vp)],
            "cs_batch_evaluate_timestep": [vp],
            "cs_batch_advance": [vp, i32, i32],
            "cs_batch_time_steps": [vp, i32, ctypes.POINTER(ctypes.c_float)],
            "cs_batch_autotune": [vp, i32, ctypes.POINTER(i32)],
            "cs_kernel_units": [vp, i32, ctypes.POINTER(i64)],
            "cs_last_error": [ctypes.c_char_p, ctypes.c_size_t],
            "cs_device_info": [ctypes.c_char_p, ctypes.c_size_t],
            "cs_dump_write": [ctypes.c_char_p, cfgp, ctypes.POINTER(DumpMeta), vp, vp, vp, vp, vp],
            "cs_dump_read": [ctypes.c_char_p, cfgp, ctypes.POINTER(DumpMeta), vp, vp, vp, vp, vp],
            "cs_dump_info": [ctypes.c_char_p, ctypes.POINTER(DumpMeta)],
            "cs_write_dump": [vp, ctypes.c_char_p, ctypes.POINTER(DumpMeta)],
            "cs_read_dump": [vp, ctypes.c_char_p, ctypes.POINTER(DumpMeta)],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        _lib = L
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(512)
    lib().cs_last_error(buf, 512)
    return buf.value.decode()


def _check(rc, what):
    if rc != 0:
        raise CsError(f"{what} failed ({rc}): {last_error()}", rc)


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def sizes(cfg):
    s = Sizes()
    _check(lib().cs_sizes_of(ctypes.byref(cfg), ctypes.byref(s)), "cs_sizes_of")
    return s


class HostState:
    """Host arrays in the reference patch layout (Fortran order, 6 patches)."""

    def __init__(self, cfg):
        s = sizes(cfg)
        self.cfg = cfg
        self.U = np.zeros(s.n_cell * 8)
        self.ent = np.zeros(s.n_cell)
        self.Bx = np.zeros(s.n_face)
        self.By = np.zeros(s.n_face)
        self.Bz = np.zeros(s.n_face)

    def arrays(self):
        return [self.U, self.ent, self.Bx, self.By, self.Bz]

    def copy(self):
        o = HostState.__new__(HostState)
        o.cfg = self.cfg
        o.U, o.ent, o.Bx, o.By, o.Bz = (a.copy() for a in self.arrays())
        return o

    # views with named axes: U[s, var, k, j, i], faces[s, k, j, i]
    def views(self):
        c = self.cfg
        nc, nrc = c.ntp + 2 * c.nghost, c.nr + 2 * c.nghost
        nf, nrf = nc + 1, nrc + 1
        return (self.U.reshape(6, 8, nrc, nc, nc), self.ent.reshape(6, nrc, nc, nc),
                self.Bx.reshape(6, nrf, nf, nf), self.By.reshape(6, nrf, nf, nf),
                self.Bz.reshape(6, nrf, nf, nf))

    def interior(self):
        """U (6,8,nr,N,N) and ent of the owned zones."""
        c = self.cfg
        g, N, NR = c.nghost, c.ntp, c.nr
        U, ent = self.views()[:2]
        return U[:, :, g:g + NR, g:g + N, g:g + N], ent[:, g:g + NR, g:g + N, g:g + N]


class DumpMeta(ctypes.Structure):
    """cs_dump_meta: the header fields of a checkpoint (riemann.in's time,
    dtcur, n_cc_components and dump number, the step count, the format
    version and the configuration the state was written with)."""
    _fields_ = [("time", ctypes.c_double), ("dtcur", ctypes.c_double), ("filenum", ctypes.c_int64),
                ("step", ctypes.c_int64), ("n_cc_components", ctypes.c_int32), ("version", ctypes.c_int32),
                ("cfg", Config)]


def _path(path):
    return os.fsencode(os.fspath(path))


def write_dump(path, st, time=0.0, dtcur=0.0, filenum=0, step=0):
    """Checkpoint of a host state (cs_dump_write): the riemann.in fields and
    the five arrays in one versioned, checksummed file."""
    m = DumpMeta(float(time), float(dtcur), int(filenum), int(step), 8, 0)
    _check(lib().cs_dump_write(_path(path), ctypes.byref(st.cfg), ctypes.byref(m), *[_p(a) for a in st.arrays()]),
           "cs_dump_write")


def read_dump(path, cfg):
    """(HostState, DumpMeta) of a checkpoint written on cfg's grid (cs_dump_read)."""
    st = HostState(cfg)
    m = DumpMeta()
    _check(lib().cs_dump_read(_path(path), ctypes.byref(cfg), ctypes.byref(m), *[_p(a) for a in st.arrays()]),
           "cs_dump_read")
    return st, m


def dump_info(path):
    """The checked header of a checkpoint (cs_dump_info)."""
    m = DumpMeta()
    _check(lib().cs_dump_info(_path(path), ctypes.byref(m)), "cs_dump_info")
    return m


def init_problem(cfg, problem, seed=1234):
    st = HostState(cfg)
    _check(lib().cs_init_problem(ctypes.byref(cfg), problem, seed, *[_p(a) for a in st.arrays()]),
           "cs_init_problem")
    return st


def metric(cfg):
    s = sizes(cfg)
    vol, xa, ya, za = (np.zeros(n) for n in (s.n_cell, s.n_face, s.n_face, s.n_face))
    _check(lib().cs_build_metric(ctypes.byref(cfg), _p(vol), _p(xa), _p(ya), _p(za)),
           "cs_build_metric")
    return vol, xa, ya, za


def unit_lattice(cfg):
    s = sizes(cfg)
    a = np.zeros(s.n_vert2d * 3)
    _check(lib().cs_build_unit_lattice(ctypes.byref(cfg), _p(a)), "cs_build_unit_lattice")
    return a


class CubeSphere:
    """One shell (all six sectors) resident on the GPU.

    Methods mirror the reference's per-patch call chain:
      halo_exchange()        FETCH_BUFFERS_FILL_IN_UX/UY/UY_TOP/UZ (330-333)
      make_facial_fluxes()   LIMIT_ALL_SLOPES_P1 + MAKE_FACIAL_FLUXES (410, 450)
      make_edge_efields()    MAKE_EDGE_EFIELDS_P1 (460)
      update_stage(irk)      UPDATE_IMEX_RK_STEP_P1 + CLEANUP_AFTER_UPDATE (484, 485)
      evaluate_timestep()    EVALUATE_TIMESTEP (490)
    """

    def __init__(self, cfg, _borrowed=None):
        self.cfg = cfg
        self.owner = _borrowed is None
        if _borrowed is None:
            h = ctypes.c_void_p()
            _check(lib().cs_create(ctypes.byref(cfg), ctypes.byref(h)), "cs_create")
        else:
            h = _borrowed
        self.h = h
        self.sizes = sizes(cfg)

    def close(self):
        if self.h and self.owner:
            lib().cs_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, st):
        _check(lib().cs_upload_state(self.h, *[_p(a) for a in st.arrays()]), "cs_upload_state")

    def download(self):
        st = HostState(self.cfg)
        _check(lib().cs_download_state(self.h, *[_p(a) for a in st.arrays()]), "cs_download_state")
        return st

    def write_dump(self, path, time=0.0, filenum=0, step=0):
        """Checkpoint of U^n straight from the device (cs_write_dump); dtcur
        is the handle's current time step."""
        m = DumpMeta(float(time), 0.0, int(filenum), int(step), 8, 0)
        _check(lib().cs_write_dump(self.h, _path(path), ctypes.byref(m)), "cs_write_dump")

    def read_dump(self, path):
        """Restart from a checkpoint (cs_read_dump): the state uploaded as by
        upload() and the time step set to the dump's dtcur; returns its
        DumpMeta."""
        m = DumpMeta()
        _check(lib().cs_read_dump(self.h, _path(path), ctypes.byref(m)), "cs_read_dump")
        return m

    def halo_exchange(self):
        _check(lib().cs_halo_exchange(self.h), "cs_halo_exchange")

    def halo_exchange_p1(self, patches, stream=None):
        """cs_halo_exchange_p1 over six P1 patches (dicts of CUDA tensors with
        U, ent, Bxf, Byf, Bzf in the layouts of cs_p1_patch)."""
        lists = []
        for k in ("U", "ent", "Bxf", "Byf", "Bzf"):
            arr = (ctypes.c_void_p * 6)(*[p[k].data_ptr() for p in patches])
            lists.append(arr)
        st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        _check(lib().cs_halo_exchange_p1(self.h, *[ctypes.cast(a, ctypes.c_void_p) for a in lists], st),
               "cs_halo_exchange_p1")

    def make_facial_fluxes(self):
        _check(lib().cs_make_facial_fluxes(self.h), "cs_make_facial_fluxes")

    def make_edge_efields(self):
        _check(lib().cs_make_edge_efields(self.h), "cs_make_edge_efields")

    def update_stage(self, irk):
        _check(lib().cs_update_stage(self.h, irk), "cs_update_stage")

    def evaluate_timestep(self):
        _check(lib().cs_evaluate_timestep(self.h), "cs_evaluate_timestep")

    def rk_stage(self, irk):
        _check(lib().cs_rk_stage(self.h, irk), "cs_rk_stage")

    def advance(self, nsteps, use_graph=True):
        _check(lib().cs_advance(self.h, nsteps, 1 if use_graph else 0), "cs_advance")

    @property
    def dt(self):
        v = ctypes.c_double()
        _check(lib().cs_get_dt(self.h, ctypes.byref(v)), "cs_get_dt")
        return v.value

    @dt.setter
    def dt(self, value):
        _check(lib().cs_set_dt(self.h, float(value)), "cs_set_dt")

    def fluxes(self, d):
        F = np.zeros(self.sizes.n_face * 9)
        S = np.zeros(self.sizes.n_face * 2)
        _check(lib().cs_download_fluxes(self.h, d, _p(F), _p(S)), "cs_download_fluxes")
        return F, S

    def emf(self):
        E = [np.zeros(self.sizes.n_face) for _ in range(3)]
        _check(lib().cs_download_emf(self.h, *[_p(a) for a in E]), "cs_download_emf")
        return E

    def conservation_sums(self):
        out = np.zeros(9)
        _check(lib().cs_conservation_sums(self.h, _p(out)), "cs_conservation_sums")
        return out

    def max_divb(self):
        v = ctypes.c_double()
        _check(lib().cs_max_divb(self.h, ctypes.byref(v)), "cs_max_divb")
        return v.value

    def time_steps(self, nsteps, use_graph=True):
        ms = ctypes.c_float()
        _check(lib().cs_time_steps(self.h, nsteps, 1 if use_graph else 0, ctypes.byref(ms)),
               "cs_time_steps")
        return ms.value

    def time_kernel(self, kernel, iters=10):
        """Average duration (ms) of one kernel measured with hipEvents."""
        ms = ctypes.c_float()
        _check(lib().cs_time_kernel(self.h, kernel, iters, ctypes.byref(ms)), "cs_time_kernel")
        return ms.value

    def profile_steps(self, nsteps):
        """Run nsteps steps launch by launch with an event pair around every
        launch; returns {kernel id: (average ms per launch, launches)}."""
        ms = (ctypes.c_float * len(KERNEL_NAMES))()
        calls = (ctypes.c_int32 * len(KERNEL_NAMES))()
        _check(lib().cs_profile_steps(self.h, nsteps, ms, calls), "cs_profile_steps")
        return {k: (ms[k] / calls[k], calls[k]) for k in range(len(KERNEL_NAMES)) if calls[k]}

    def init_problem_p1(self, problem, seed=0, time=0.0):
        """INIT_PATCH_CUBE_SPHERE with PROBLEM_GENERATOR_U_A on the device
        (recon RECON_P1ADP): problems 0, 2 (the reference's), 102, 103."""
        _check(lib().cs_init_problem_p1(self.h, int(problem), int(seed), float(time)), "cs_init_problem_p1")

    def add_boundary_p1(self, inner, outer, problem, seed=0, time=0.0):
        """ADD_BOUNDARY: the problem's state on the inner and / or outer
        radial ghost layers."""
        _check(lib().cs_add_boundary_p1(self.h, int(bool(inner)), int(bool(outer)), int(problem), int(seed),
                                        float(time)), "cs_add_boundary_p1")

    def profile_steps_p1(self, nsteps):
        """The same for a handle of recon RECON_P1ADP: {P1_KERNEL_NAMES entry:
        (average ms per launch group, launch groups)}."""
        n = len(P1_KERNEL_NAMES)
        ms = (ctypes.c_float * n)()
        calls = (ctypes.c_int32 * n)()
        _check(lib().cs_profile_steps_p1(self.h, nsteps, ms, calls), "cs_profile_steps_p1")
        return {P1_KERNEL_NAMES[k]: (ms[k] / calls[k], calls[k]) for k in range(n) if calls[k]}

    def autotune(self, iters=3):
        """Pick each stage launch's workgroup size by timing; returns {kernel name: size}."""
        bs = (ctypes.c_int32 * len(KERNEL_NAMES))()
        _check(lib().cs_autotune(self.h, iters, bs), "cs_autotune")
        return {KERNEL_NAMES[k]: bs[k] for k in range(len(KERNEL_NAMES))}

    def kernel_units(self, kernel):
        n = ctypes.c_int64()
        _check(lib().cs_kernel_units(self.h, kernel, ctypes.byref(n)), "cs_kernel_units")
        return n.value


BATCH_PHASES = ["exchange", "radial_pad", "flux_xy", "flux_z", "emf", "update_faces", "update_cells",
                "implicit_source", "cfl"]


class Batch:
    """Independent shells advanced together: every phase of a stage is one
    launch over all shells and a step is one hipGraph.  The shells share
    the precision and the time integrator."""

    def __init__(self, cfgs):
        self.cfgs = list(cfgs)
        arr = (Config * len(self.cfgs))(*self.cfgs)
        b = ctypes.c_void_p()
        _check(lib().cs_batch_create(arr, len(self.cfgs), ctypes.byref(b)), "cs_batch_create")
        self.b = b
        self.shells = []
        for i, c in enumerate(self.cfgs):
            h = ctypes.c_void_p()
            _check(lib().cs_batch_shell(self.b, i, ctypes.byref(h)), "cs_batch_shell")
            self.shells.append(CubeSphere(c, _borrowed=h))

    def close(self):
        if self.b:
            for s in self.shells:
                s.close()
            lib().cs_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def evaluate_timestep(self):
        _check(lib().cs_batch_evaluate_timestep(self.b), "cs_batch_evaluate_timestep")

    def advance(self, nsteps, use_graph=True):
        _check(lib().cs_batch_advance(self.b, nsteps, 1 if use_graph else 0), "cs_batch_advance")

    def time_steps(self, nsteps):
        ms = ctypes.c_float()
        _check(lib().cs_batch_time_steps(self.b, nsteps, ctypes.byref(ms)), "cs_batch_time_steps")
        return ms.value

    def autotune(self, iters=3):
        """Pick each phase's workgroup size by timing; returns {phase: size}."""
        bs = (ctypes.c_int32 * len(BATCH_PHASES))()
        _check(lib().cs_batch_autotune(self.b, iters, bs), "cs_batch_autotune")
        return dict(zip(BATCH_PHASES, list(bs)))


def device_info():
    buf = ctypes.create_string_buffer(256)
    rc = lib().cs_device_info(buf, 256)
    return buf.value.decode() if rc == 0 else None


# ------------------------------------------------------------------ P1 geometry
GEOMETRY_FACE_KEYS = ("pts", "wts", "normal", "area", "centroid")


class FaceQuad(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in GEOMETRY_FACE_KEYS]


class Geometry(ctypes.Structure):
    _fields_ = [("vert", ctypes.c_void_p), ("mask", ctypes.c_void_p), ("xf", FaceQuad), ("yf", FaceQuad),
                ("zf", FaceQuad), ("scale", ctypes.c_void_p), ("xb", FaceQuad), ("yb", FaceQuad),
                ("zb", FaceQuad), ("v_pts", ctypes.c_void_p),
                ("v_wts", ctypes.c_void_p), ("v_tan", ctypes.c_void_p), ("v_vol", ctypes.c_void_p),
                ("v_cen", ctypes.c_void_p), ("moments", ctypes.c_void_p)]


def geometry_sizes(cfg):
    """Element counts over the six sectors: vertices, zones, x-, y-, z-faces."""
    n = np.zeros(5, np.int64)
    _check(lib().cs_geometry_sizes(ctypes.byref(cfg), _p(n)), "cs_geometry_sizes")
    return dict(zip(("vert", "zone", "xf", "yf", "zf"), (int(x) for x in n)))


def build_geometry(cfg, download=True, skip=()):
    """The quadratic isoparametric geometry of the six sectors built on the GPU
    (EXTRACT_VERT_COORDS_PATCH, COORDS_TO_WT_QUADPTS_FACES_ISO2,
    COORDS_TO_WT_QUADPTS_VOLUME_ISO2, INIT_BASIS_FUNCTION_COEFS).  Returns
    (arrays, kernel milliseconds); with download=False only the times.  Zone
    arrays named in skip (e.g. "v_tan", "v_pts") are not copied back."""
    z = geometry_sizes(cfg)
    ms = (ctypes.c_float * 4)()
    if not download:
        _check(lib().cs_build_geometry(ctypes.byref(cfg), None, ms), "cs_build_geometry")
        return None, list(ms)
    per = dict(vert=3 * z["vert"], scale=z["zone"], v_pts=24 * z["zone"], v_wts=8 * z["zone"],
               v_tan=72 * z["zone"], v_vol=z["zone"], v_cen=3 * z["zone"], moments=4 * z["zone"])
    a = {k: np.zeros(n) for k, n in per.items() if k not in skip}
    a["mask"] = np.zeros(z["zone"], np.int32)
    G = Geometry()
    for k in ("vert", "mask", "scale", "v_pts", "v_wts", "v_tan", "v_vol", "v_cen", "moments"):
        setattr(G, k, a[k].ctypes.data if k in a else None)
    for f in ("xf", "yf", "zf"):
        n = z[f]
        a[f] = dict(pts=np.zeros(15 * n), wts=np.zeros(5 * n), normal=np.zeros((15 if f == "zf" else 3) * n),
                    area=np.zeros(n), centroid=np.zeros(3 * n))
        fq = getattr(G, f)
        for k in GEOMETRY_FACE_KEYS:
            setattr(fq, k, a[f][k].ctypes.data)
        bname = f[0] + "b"   # B-field points of the same faces (no area, centroid)
        a[bname] = dict(pts=np.zeros(15 * n), wts=np.zeros(5 * n), normal=np.zeros((15 if f == "zf" else 3) * n))
        bq = getattr(G, bname)
        for k in GEOMETRY_FACE_KEYS[:3]:
            setattr(bq, k, a[bname][k].ctypes.data)
    _check(lib().cs_build_geometry(ctypes.byref(cfg), ctypes.byref(G), ms), "cs_build_geometry")
    return a, list(ms)


def mode_indices(order):
    """BUILD_MODE_INDICES: the monomial powers (i, j, k) of each mode."""
    ii, jj, kk = (np.zeros(20, np.int32) for _ in range(3))
    n = ctypes.c_int32()
    _check(lib().cs_mode_indices(order, ctypes.byref(n), _p(ii), _p(jj), _p(kk)), "cs_mode_indices")
    return [(int(ii[m]), int(jj[m]), int(kk[m])) for m in range(n.value)]


def stencil_indexing(cfg):
    """STENCIL_INDEXING_TRICORNERS: int32 (12, 2, N+2g, N+2g) view of the
    reference's istencil(1-g:N+g, 1-g:N+g, 2, 12): [dir-1, (x, y), y, x]."""
    nc = cfg.ntp + 2 * cfg.nghost
    a = np.zeros(nc * nc * 24, np.int32)
    _check(lib().cs_stencil_indexing(ctypes.byref(cfg), _p(a)), "cs_stencil_indexing")
    return a


# ---------------------------------------------------------------------------
# LIMIT_ALL_SLOPES_P1 of one patch (cs_limit_all_slopes_p1)

P1_SCALARS = ("gamma", "molewt", "rhofloor", "prsfloor", "smallnum", "flatten_coef", "prs_flatten_coef")
P1_ARRAYS = ("vert", "mask", "stencil", "xf_normal", "xf_area", "xf_centroid", "yf_normal", "yf_area",
             "yf_centroid", "zf_area", "zf_centroid", "xb_pts", "xb_wts", "xb_normal", "yb_pts", "yb_wts",
             "yb_normal", "zb_pts", "zb_wts", "zb_normal", "scale", "v_centroid",
             "U", "ent", "Bxf", "Byf", "Bzf", "V", "B1_zc", "B2_zc", "B3_zc", "Ex_zc", "Ey_zc", "Ez_zc",
             "gamma_eos", "gamma_sound", "mean_mol_wt", "divvel", "msonic", "flatten", "work",
             "Bxf_q", "Byf_q", "Bzf_q")


class P1Patch(ctypes.Structure):
    """cs_p1_patch: the arguments of LIMIT_ALL_SLOPES_P1 for one patch."""
    _fields_ = ([("ntp", ctypes.c_int32), ("nr", ctypes.c_int32), ("nghost", ctypes.c_int32),
                 ("itri", ctypes.c_int32 * 4), ("pad", ctypes.c_int32)]
                + [(k, ctypes.c_double) for k in P1_SCALARS]
                + [(k, ctypes.c_void_p) for k in P1_ARRAYS]
                + [("V_min_local", ctypes.c_void_p), ("V_max_local", ctypes.c_void_p)])


def p1_counts(N, NR, g):
    """Element count of every array of a patch, from the reference's layouts."""
    nz, nzr = N + 2 * g, NR + 2 * g
    nf, nfr = nz + 1, nzr + 1
    zone, xf, zf, fc = nz * nz * nzr, nf * nz * nzr, nz * nz * nfr, nf * nf * nfr
    c = dict(vert=3 * fc, mask=zone, stencil=nz * nz * 24, xf_normal=3 * xf, xf_area=xf, xf_centroid=3 * xf,
             yf_normal=3 * xf, yf_area=xf, yf_centroid=3 * xf, zf_area=zf, zf_centroid=3 * zf,
             xb_pts=15 * xf, xb_wts=5 * xf, xb_normal=3 * xf, yb_pts=15 * xf, yb_wts=5 * xf, yb_normal=3 * xf,
             zb_pts=15 * zf, zb_wts=5 * zf, zb_normal=15 * zf, scale=zone, v_centroid=3 * zone,
             U=32 * zone, ent=4 * zone, Bxf=fc, Byf=fc, Bzf=fc, V=32 * zone)
    for k in ("B1_zc", "B2_zc", "B3_zc", "Ex_zc", "Ey_zc", "Ez_zc"):
        c[k] = 4 * zone
    for k in ("gamma_eos", "gamma_sound", "mean_mol_wt", "divvel", "msonic", "flatten", "work"):
        c[k] = zone
    for k in ("Bxf_q", "Byf_q", "Bzf_q"):
        c[k] = 5 * fc
    return c


P1_STAGES = ("prim", "face_zero+tricorner", "flatten", "weno+positivity", "face_pts", "clsq+E", "weno_E",
             "minmax")


def limit_all_slopes_p1(N, NR, g, scalars, arrays, itri=(1, 1, 1, 1), stream=None, minmax=True, timed=False,
                        noalloc=False):
    """Run LIMIT_ALL_SLOPES_P1 on one patch whose arrays are torch tensors on
    the GPU (float64, int32 for mask and stencil), contiguous, in the
    reference's layouts; they are updated in place.  Every tensor's size is
    checked against the layout before the launch.  Returns (V_min_local,
    V_max_local) as numpy arrays of 9 when minmax, and with timed=True also
    the per-stage milliseconds (P1_STAGES).  noalloc=True calls
    cs_limit_all_slopes_p1_noalloc (work memory in the library's fixed
    per-GPU array, nothing allocated)."""
    import torch
    want = p1_counts(N, NR, g)
    s = P1Patch()
    s.ntp, s.nr, s.nghost = N, NR, g
    for q in range(4):
        s.itri[q] = int(itri[q])
    for k in P1_SCALARS:
        setattr(s, k, float(scalars[k]))
    for k in P1_ARRAYS:
        t = arrays[k]
        dt = torch.int32 if k in ("mask", "stencil") else torch.float64
        if not (t.is_cuda and t.dtype == dt and t.is_contiguous() and t.numel() == want[k]):
            raise CsError(f"limit_all_slopes_p1: {k} must be a contiguous {dt} CUDA tensor of {want[k]} "
                          f"elements (got {t.dtype}, {t.numel()}, cuda={t.is_cuda})")
        setattr(s, k, t.data_ptr())
    vmin = np.zeros(9)
    vmax = np.zeros(9)
    if minmax:
        s.V_min_local = vmin.ctypes.data
        s.V_max_local = vmax.ctypes.data
    st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
    if timed:
        ms = (ctypes.c_float * len(P1_STAGES))()
        _check(lib().cs_limit_all_slopes_p1_timed(ctypes.byref(s), st, ms), "cs_limit_all_slopes_p1_timed")
        return vmin, vmax, list(ms)
    if noalloc:
        _check(lib().cs_limit_all_slopes_p1_noalloc(ctypes.byref(s), st), "cs_limit_all_slopes_p1_noalloc")
    else:
        _check(lib().cs_limit_all_slopes_p1(ctypes.byref(s), st), "cs_limit_all_slopes_p1")
    return vmin, vmax


# ---------------------------------------------------------------------------
# Third order: the NUMORDER 3 volume geometry and LIMIT_ALL_SLOPES_P12
# (cs_build_volume_p12, cs_limit_all_slopes_p12)

class VolumeP12(ctypes.Structure):
    """cs_volume_p12: host outputs of the third-order volume geometry."""
    _fields_ = [(k, ctypes.c_void_p) for k in ("v_pts", "v_wts", "v_vol", "v_cen", "moments")]


def build_volume_p12(cfg, skip=()):
    """COORDS_TO_WT_QUADPTS_VOLUME_ISO2 with the 27-point rule and the ten basis
    moments of INIT_BASIS_FUNCTION_COEFS at third order, for the six sectors,
    built on the GPU: dict of v_pts (zone, 3, 27), v_wts (zone, 27), v_vol,
    v_cen (zone, 3), moments (zone, 10), sector slowest; arrays named in skip
    are not copied back."""
    nz = geometry_sizes(cfg)["zone"]
    per = dict(v_pts=81, v_wts=27, v_vol=1, v_cen=3, moments=10)
    a = {k: np.zeros(m * nz) for k, m in per.items() if k not in skip}
    o = VolumeP12()
    for k in per:
        setattr(o, k, a[k].ctypes.data if k in a else None)
    _check(lib().cs_build_volume_p12(ctypes.byref(cfg), ctypes.byref(o)), "cs_build_volume_p12")
    return a


class P12Patch(ctypes.Structure):
    """cs_p12_patch: a cs_p1_patch whose U and V hold ten modes, plus the ten
    basis moments."""
    _fields_ = [("p1", P1Patch), ("moments", ctypes.c_void_p)]


def limit_all_slopes_p12(N, NR, g, scalars, arrays, moments, itri=(1, 1, 1, 1), stream=None, timed=False):
    """LIMIT_ALL_SLOPES_P12 on one patch, the third-order WENO-AO
    reconstruction (the P1 steps up to the WENO, the hybridisation with the
    quadratic central stencil, the positivity limiter and the face-point B
    with the quadratic modes, the 30-unknown constrained least squares of the
    zone B, the zone E and its WENO-AO, the min/max diagnostics), in place on
    torch CUDA tensors in the reference's layouts: as limit_all_slopes_p1,
    with U and V of ten modes (zone, 8, 10), B1..3_zc and Ex..Ez_zc of ten,
    v_centroid from the 27-point rule, and moments (zone, 10) a float64 CUDA
    tensor.  Returns (V_min_local, V_max_local) as numpy arrays of 9, and with
    timed=True also the per-stage milliseconds (P1_STAGES)."""
    import torch
    want = p1_counts(N, NR, g)
    nzone = want["mask"]
    want["U"] = want["V"] = 80 * nzone
    for k in ("B1_zc", "B2_zc", "B3_zc", "Ex_zc", "Ey_zc", "Ez_zc"):
        want[k] = 10 * nzone
    s = P12Patch()
    q = s.p1
    q.ntp, q.nr, q.nghost = N, NR, g
    for i in range(4):
        q.itri[i] = int(itri[i])
    for k in P1_SCALARS:
        setattr(q, k, float(scalars[k]))
    tensors = dict(arrays, moments=moments)
    want["moments"] = 10 * nzone
    for k in P1_ARRAYS + ("moments",):
        t = tensors[k]
        dt = torch.int32 if k in ("mask", "stencil") else torch.float64
        if not (t.is_cuda and t.dtype == dt and t.is_contiguous() and t.numel() >= want[k]):
            raise CsError(f"limit_all_slopes_p12: {k} must be a contiguous {dt} CUDA tensor of at least "
                          f"{want[k]} elements (got {t.dtype}, {t.numel()}, cuda={t.is_cuda})")
        if k == "moments":
            s.moments = t.data_ptr()
        else:
            setattr(q, k, t.data_ptr())
    vmin, vmax = np.zeros(9), np.zeros(9)
    q.V_min_local, q.V_max_local = vmin.ctypes.data, vmax.ctypes.data
    st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
    if timed:
        ms = (ctypes.c_float * len(P1_STAGES))()
        _check(lib().cs_limit_all_slopes_p12_timed(ctypes.byref(s), st, ms), "cs_limit_all_slopes_p12_timed")
        return vmin, vmax, list(ms)
    _check(lib().cs_limit_all_slopes_p12(ctypes.byref(s), st), "cs_limit_all_slopes_p12")
    return vmin, vmax


# ---------------------------------------------------------------------------
# MAKE_FACIAL_FLUXES of one patch (cs_make_facial_fluxes_p1)

P1F_SCALARS = ("gamma", "rhofloor", "prsfloor", "smallnum")
P1F_IN = ("vert", "mask", "stencil", "xf_pts", "xf_normal", "xf_area", "xf_centroid", "yf_pts", "yf_normal",
          "yf_area", "yf_centroid", "zf_pts", "zf_normal", "zf_area", "zf_centroid", "scale", "v_volume",
          "v_centroid", "V", "ent", "flatten")
P1F_OUT = ("S_L", "S_R", "S_D", "S_U", "S_B", "S_T", "Fluct_x_Pl", "Fluct_y_Pl", "Fluct_z_Pl", "Fluct_x_Mi",
           "Fluct_y_Mi", "Fluct_z_Mi", "dU_dt", "EntFlux_x", "EntFlux_y", "EntFlux_z", "Efx", "Efy", "Efz")


class P1Flux(ctypes.Structure):
    """cs_p1_flux: the arguments of MAKE_FACIAL_FLUXES for one patch."""
    _fields_ = ([("ntp", ctypes.c_int32), ("nr", ctypes.c_int32), ("nghost", ctypes.c_int32),
                 ("pad", ctypes.c_int32)]
                + [(k, ctypes.c_double) for k in P1F_SCALARS]
                + [(k, ctypes.c_void_p) for k in P1F_IN + P1F_OUT])


def p1f_counts(N, NR, g):
    c = p1_counts(N, NR, g)
    zone, fc = c["mask"], c["Bxf"]
    xf, zf = c["xf_area"], c["zf_area"]
    c.update(xf_pts=15 * xf, yf_pts=15 * xf, zf_pts=15 * zf, zf_normal=15 * zf, v_volume=zone, dU_dt=8 * zone)
    for k in ("S_L", "S_R", "S_D", "S_U", "S_B", "S_T", "EntFlux_x", "EntFlux_y", "EntFlux_z", "Efx", "Efy", "Efz"):
        c[k] = fc
    for k in ("Fluct_x_Pl", "Fluct_y_Pl", "Fluct_z_Pl", "Fluct_x_Mi", "Fluct_y_Mi", "Fluct_z_Mi"):
        c[k] = 8 * fc
    return c


def make_facial_fluxes_p1(N, NR, g, scalars, arrays, stream=None):
    """MAKE_FACIAL_FLUXES on one patch whose arrays are contiguous torch
    tensors on the GPU in the reference's layouts (float64; int32 for mask and
    stencil); outputs are written in place.  Sizes are checked first."""
    import torch
    want = p1f_counts(N, NR, g)
    s = P1Flux()
    s.ntp, s.nr, s.nghost = N, NR, g
    for k in P1F_SCALARS:
        setattr(s, k, float(scalars[k]))
    for k in P1F_IN + P1F_OUT:
        t = arrays[k]
        dt = torch.int32 if k in ("mask", "stencil") else torch.float64
        if not (t.is_cuda and t.dtype == dt and t.is_contiguous() and t.numel() == want[k]):
            raise CsError(f"make_facial_fluxes_p1: {k} must be a contiguous {dt} CUDA tensor of {want[k]} "
                          f"elements (got {t.dtype}, {t.numel()}, cuda={t.is_cuda})")
        setattr(s, k, t.data_ptr())
    st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
    _check(lib().cs_make_facial_fluxes_p1(ctypes.byref(s), st), "cs_make_facial_fluxes_p1")


# ---------------------------------------------------------------------------
# UPDATE_IMEX_RK_STEP_P1 of one patch (cs_update_imex_rk_step_p1)

P1U_SCALARS = ("dtmove", "gamma", "rhofloor", "prsfloor")
P1U_ARRAYS = ("vert", "xf_area", "yf_area", "zf_area", "v_volume", "U", "ent", "Bxf", "Byf", "Bzf", "dU_dt",
              "EntFlux_x", "EntFlux_y", "EntFlux_z", "Efx", "Efy", "Efz", "flag_entprs")


class P1Update(ctypes.Structure):
    """cs_p1_update: the arguments of UPDATE_IMEX_RK_STEP_P1 for one patch."""
    _fields_ = ([("ntp", ctypes.c_int32), ("nr", ctypes.c_int32), ("nghost", ctypes.c_int32),
                 ("pad", ctypes.c_int32)]
                + [(k, ctypes.c_double) for k in P1U_SCALARS]
                + [(k, ctypes.c_void_p) for k in P1U_ARRAYS])


def update_imex_rk_step_p1(N, NR, g, scalars, arrays, stream=None):
    """UPDATE_IMEX_RK_STEP_P1 on one patch of contiguous float64 torch tensors
    on the GPU in the reference's layouts (sizes checked first)."""
    import torch
    want = p1f_counts(N, NR, g)
    want["flag_entprs"] = want["mask"]
    s = P1Update()
    s.ntp, s.nr, s.nghost = N, NR, g
    for k in P1U_SCALARS:
        setattr(s, k, float(scalars[k]))
    for k in P1U_ARRAYS:
        t = arrays[k]
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() == want[k]):
            raise CsError(f"update_imex_rk_step_p1: {k} must be a contiguous float64 CUDA tensor of {want[k]} "
                          f"elements (got {t.dtype}, {t.numel()}, cuda={t.is_cuda})")
        setattr(s, k, t.data_ptr())
    st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
    _check(lib().cs_update_imex_rk_step_p1(ctypes.byref(s), st), "cs_update_imex_rk_step_p1")


# ---------------------------------------------------------------------------
# CLEANUP_AFTER_UPDATE and EVALUATE_TIMESTEP of one patch, on the fields of
# cs_p1_patch (cs_cleanup_after_update_p1, cs_evaluate_timestep_p1)

P1C_ARRAYS = ("mask", "stencil", "xf_normal", "yf_normal", "zf_centroid", "scale", "U", "ent", "Bxf", "Byf", "Bzf",
              "V", "B1_zc", "B2_zc", "B3_zc", "gamma_eos", "gamma_sound", "mean_mol_wt")
P1T_ARRAYS = ("mask", "scale", "U", "ent", "V", "gamma_sound", "msonic")


def _p1_subset(who, N, NR, g, scalars, arrays, names, itri=(1, 1, 1, 1)):
    import torch
    want = p1_counts(N, NR, g)
    s = P1Patch()
    s.ntp, s.nr, s.nghost = N, NR, g
    for q in range(4):
        s.itri[q] = int(itri[q])
    for k in P1_SCALARS:
        setattr(s, k, float(scalars[k]))
    for k in names:
        t = arrays[k]
        dt = torch.int32 if k in ("mask", "stencil") else torch.float64
        if not (t.is_cuda and t.dtype == dt and t.is_contiguous() and t.numel() == want[k]):
            raise CsError(f"{who}: {k} must be a contiguous {dt} CUDA tensor of {want[k]} "
                          f"elements (got {t.dtype}, {t.numel()}, cuda={t.is_cuda})")
        setattr(s, k, t.data_ptr())
    return s


def cleanup_after_update_p1(N, NR, g, scalars, arrays, itri=(1, 1, 1, 1), stream=None):
    """CLEANUP_AFTER_UPDATE on one patch of contiguous torch tensors on the GPU
    in the reference's layouts (P1C_ARRAYS are read; sizes checked first)."""
    s = _p1_subset("cleanup_after_update_p1", N, NR, g, scalars, arrays, P1C_ARRAYS, itri)
    st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
    _check(lib().cs_cleanup_after_update_p1(ctypes.byref(s), st), "cs_cleanup_after_update_p1")


def evaluate_timestep_p1(N, NR, g, scalars, arrays, cfl, dtcur, out=None, stream=None):
    """EVALUATE_TIMESTEP on one patch: returns the new time step as a float64
    CUDA tensor of one element (out, when given, receives it) without waiting
    on the stream."""
    import torch
    s = _p1_subset("evaluate_timestep_p1", N, NR, g, scalars, arrays, P1T_ARRAYS)
    if out is None:
        out = torch.empty(1, dtype=torch.float64, device=arrays["U"].device)
    if not (out.is_cuda and out.dtype == torch.float64 and out.numel() >= 1):
        raise CsError("evaluate_timestep_p1: out must be a float64 CUDA tensor")
    st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
    _check(lib().cs_evaluate_timestep_p1(ctypes.byref(s), float(cfl), float(dtcur), ctypes.c_void_p(out.data_ptr()),
                                         st), "cs_evaluate_timestep_p1")
    return out


# ---------------------------------------------------------------------------
# MAKE_EDGE_EFIELDS_P1 of one patch (cs_make_edge_efields_p1)

P1E_IN = ("vert", "mask", "Ex_zc", "Ey_zc", "Ez_zc", "Bxf_q", "Byf_q", "Bzf_q", "S_L", "S_R", "S_D", "S_U", "S_B",
          "S_T")
P1E_OUT = ("Efx", "Efy", "Efz")


class P1Edge(ctypes.Structure):
    """cs_p1_edge: the arguments of MAKE_EDGE_EFIELDS_P1 that it reads at
    second order, for one patch."""
    _fields_ = ([("ntp", ctypes.c_int32), ("nr", ctypes.c_int32), ("nghost", ctypes.c_int32),
                 ("pad", ctypes.c_int32), ("itri", ctypes.c_int32 * 4)]
                + [(k, ctypes.c_void_p) for k in P1E_IN + P1E_OUT])


def make_edge_efields_p1(N, NR, g, arrays, itri=(1, 1, 1, 1), stream=None):
    """MAKE_EDGE_EFIELDS_P1 on one patch of contiguous torch tensors on the GPU
    in the reference's layouts (float64; int32 for mask); Efx, Efy, Efz are
    written.  Sizes are checked first."""
    import torch
    want = p1f_counts(N, NR, g)
    s = P1Edge()
    s.ntp, s.nr, s.nghost = N, NR, g
    for q in range(4):
        s.itri[q] = int(itri[q])
    for k in P1E_IN + P1E_OUT:
        t = arrays[k]
        dt = torch.int32 if k == "mask" else torch.float64
        if not (t.is_cuda and t.dtype == dt and t.is_contiguous() and t.numel() == want[k]):
            raise CsError(f"make_edge_efields_p1: {k} must be a contiguous {dt} CUDA tensor of {want[k]} "
                          f"elements (got {t.dtype}, {t.numel()}, cuda={t.is_cuda})")
        setattr(s, k, t.data_ptr())
    st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
    _check(lib().cs_make_edge_efields_p1(ctypes.byref(s), st), "cs_make_edge_efields_p1")


# ---------------------------------------------------------------------------
# BUILD_NON_STIFF_SOURCES of one patch (cs_build_non_stiff_sources_p1)

class P1Source(ctypes.Structure):
    """cs_p1_source: the arguments of BUILD_NON_STIFF_SOURCES for one patch."""
    _fields_ = [("ntp", ctypes.c_int32), ("nr", ctypes.c_int32), ("nghost", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("newtonsc", ctypes.c_double), ("mstar", ctypes.c_double),
                ("smallnum", ctypes.c_double), ("omega", ctypes.c_double * 3), ("mask", ctypes.c_void_p),
                ("v_centroid", ctypes.c_void_p), ("V", ctypes.c_void_p), ("Source", ctypes.c_void_p),
                ("dU_dt", ctypes.c_void_p)]


def build_non_stiff_sources_p1(N, NR, g, newtonsc, mstar, smallnum, omega, arrays, stream=None):
    """BUILD_NON_STIFF_SOURCES on one patch of contiguous torch tensors on the
    GPU (mask int32; v_centroid, V, Source, dU_dt float64); Source is written
    and dU_dt updated in place.  Sizes are checked first."""
    import torch
    want = p1_counts(N, NR, g)
    want.update(Source=8 * want["mask"], dU_dt=8 * want["mask"])
    s = P1Source()
    s.ntp, s.nr, s.nghost = N, NR, g
    s.newtonsc, s.mstar, s.smallnum = float(newtonsc), float(mstar), float(smallnum)
    for q in range(3):
        s.omega[q] = float(omega[q])
    for k in ("mask", "v_centroid", "V", "Source", "dU_dt"):
        t = arrays[k]
        dt = torch.int32 if k == "mask" else torch.float64
        if not (t.is_cuda and t.dtype == dt and t.is_contiguous() and t.numel() == want[k]):
            raise CsError(f"build_non_stiff_sources_p1: {k} must be a contiguous {dt} CUDA tensor of {want[k]} "
                          f"elements (got {t.dtype}, {t.numel()}, cuda={t.is_cuda})")
        setattr(s, k, t.data_ptr())
    st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
    _check(lib().cs_build_non_stiff_sources_p1(ctypes.byref(s), st), "cs_build_non_stiff_sources_p1")


# ---------------------------------------------------------------------------
# PAD_BOUNDARY_R of one patch (cs_pad_boundary_r_p1)

class P1Pad(ctypes.Structure):
    _fields_ = [("ntp", ctypes.c_int32), ("nr", ctypes.c_int32), ("nghost", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("mask", ctypes.c_void_p), ("bcarr_r", ctypes.c_void_p),
                ("xf_area", ctypes.c_void_p), ("yf_area", ctypes.c_void_p), ("zf_area", ctypes.c_void_p),
                ("v_volume", ctypes.c_void_p), ("U", ctypes.c_void_p), ("Bxf", ctypes.c_void_p),
                ("Byf", ctypes.c_void_p), ("Bzf", ctypes.c_void_p)]


def pad_boundary_r_p1(N, NR, g, arrays, stream=None):
    """PAD_BOUNDARY_R (826pad_boundary_r.f:1) on one patch of contiguous torch
    tensors on the GPU: mask and bcarr_r (1-g:N+g, 1-g:N+g, 2) int32; the face
    areas, v_volume, U (mode 1 first, any number of modes), Bxf, Byf, Bzf
    float64.  Outflow pads (bcarr_r = 2) in place.  Sizes are checked first."""
    import torch
    want = p1_counts(N, NR, g)
    nz = N + 2 * g
    want.update(bcarr_r=2 * nz * nz, v_volume=want["mask"])
    s = P1Pad()
    s.ntp, s.nr, s.nghost = N, NR, g
    for k in ("mask", "bcarr_r", "xf_area", "yf_area", "zf_area", "v_volume", "U", "Bxf", "Byf", "Bzf"):
        t = arrays[k]
        dt = torch.int32 if k in ("mask", "bcarr_r") else torch.float64
        n = t.numel()
        ok = n >= 8 * want["mask"] and n % want["mask"] == 0 if k == "U" else n == want[k]
        if not (t.is_cuda and t.dtype == dt and t.is_contiguous() and ok):
            raise CsError(f"pad_boundary_r_p1: {k} must be a contiguous {dt} CUDA tensor of "
                          f"{'8 or more modes of ' if k == 'U' else ''}{want[k]} elements "
                          f"(got {t.dtype}, {n}, cuda={t.is_cuda})")
        setattr(s, k, t.data_ptr())
    st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
    _check(lib().cs_pad_boundary_r_p1(ctypes.byref(s), st), "cs_pad_boundary_r_p1")


# ---------------------------------------------------------------------------
# Faces and edges the six P1 patches of a shell share (cs_sync_shared_p1)

FACE_PAIR = np.dtype([("pa", "<i4"), ("fa", "<i4"), ("sa", "<i4"), ("pb", "<i4"), ("fb", "<i4"), ("sb", "<i4"),
                      ("flip", "<i4"), ("pad", "<i4"), ("ia", "<i8"), ("za", "<i8"), ("aa", "<i8"), ("ib", "<i8"),
                      ("zb", "<i8"), ("ab", "<i8")])
EDGE_GROUP = np.dtype([("n", "<i4"), ("pad", "<i4"), ("p", "<i4", 3), ("f", "<i4", 3), ("sign", "<i4", 3),
                       ("pad2", "<i4"), ("idx", "<i8", 3)])


def p1_shell_pairings(geom, N, NR, g, tol=1e-9):
    """The faces on the lateral boundary of each patch (x faces i = 0, N and y
    faces j = 0, N of the owned range) paired with the same face of the
    neighbouring patch, and the lateral-boundary edges grouped with the same
    edge of the other patches (two, or three at the tri-corners), matched by
    position: geom[s] holds patch s's vert, xf_centroid, yf_centroid,
    xf_normal, yf_normal as host arrays.  Returns (faces, edges) as arrays of
    FACE_PAIR and EDGE_GROUP."""
    from scipy.spatial import cKDTree
    nz, nzr, nf, nfr = N + 2 * g, NR + 2 * g, N + 2 * g + 1, NR + 2 * g + 1
    NXF, NFC = nf * nz * nzr, nf * nf * nfr
    fci = lambda i, j, k: (i + g) + nf * ((j + g) + nf * (k + g))
    xfi = lambda i, j, k: (i + g) + nf * ((j - 1 + g) + nz * (k - 1 + g))
    yfi = lambda i, j, k: (i - 1 + g) + nz * ((j + g) + nf * (k - 1 + g))
    zi = lambda i, j, k: (i - 1 + g) + nz * ((j - 1 + g) + nz * (k - 1 + g))
    r = np.arange(1, N + 1)
    kk = np.arange(1, NR + 1)
    # faces: (patch, family, face index, zone index, area index, side, position, normal)
    rows = []
    for s, G in enumerate(geom):
        for fam, edge_vals in ((0, (0, N)), (1, (0, N))):
            cen = G["xf_centroid" if fam == 0 else "yf_centroid"].reshape(3, -1)
            nrm = G["xf_normal" if fam == 0 else "yf_normal"].reshape(3, -1)
            for b in edge_vals:
                K, T = np.meshgrid(kk, r, indexing="ij")
                K, T = K.ravel(), T.ravel()
                if fam == 0:
                    i, j = np.full_like(T, b), T
                    zone = zi(np.where(b == 0, 1, N), j, K)
                    ai = xfi(i, j, K)
                else:
                    i, j = T, np.full_like(T, b)
                    zone = zi(i, np.where(b == 0, 1, N), K)
                    ai = yfi(i, j, K)
                rows.append(dict(p=np.full_like(T, s), f=np.full_like(T, fam), idx=fci(i, j, K), zone=zone, area=ai,
                                 side=np.full_like(T, 1 if b == 0 else 0), x=cen[:, ai].T, n=nrm[:, ai].T))
    F = {k: np.concatenate([q[k] for q in rows]) for k in rows[0]}
    tree = cKDTree(F["x"])
    pairs = np.array(sorted(tree.query_pairs(tol)), dtype=np.int64).reshape(-1, 2)
    if pairs.shape[0] * 2 != F["p"].size or np.unique(pairs).size != F["p"].size:
        raise CsError("p1_shell_pairings: the patches' boundary faces do not pair up")
    faces = np.zeros(pairs.shape[0], FACE_PAIR)
    a, b = pairs[:, 0], pairs[:, 1]
    for dst, src in (("pa", "p"), ("fa", "f"), ("sa", "side"), ("ia", "idx"), ("za", "zone"), ("aa", "area")):
        faces[dst] = F[src][a]
        faces[dst[0] + "b"] = F[src][b]
    faces["flip"] = ((F["n"][a] * F["n"][b]).sum(1) < 0).astype(np.int32)
    # edges: (patch, family, index, midpoint, direction)
    rows = []
    for s, G in enumerate(geom):
        V = G["vert"].reshape(3, nfr, nf, nf)
        P = lambda i, j, k: V[:, k + g, j + g, i + g].T
        k0 = np.arange(0, NR + 1)
        for fam in range(3):
            if fam < 2:
                K, T = np.meshgrid(k0, r, indexing="ij")
                K, T = K.ravel(), T.ravel()
                for b in (0, N):
                    i, j = (T, np.full_like(T, b)) if fam == 0 else (np.full_like(T, b), T)
                    a0 = P(i - 1, j, K) if fam == 0 else P(i, j - 1, K)
                    a1 = P(i, j, K)
                    rows.append(dict(p=np.full_like(T, s), f=np.full_like(T, fam), idx=fci(i, j, K),
                                     x=0.5 * (a0 + a1), t=a1 - a0))
            else:
                ij = [(i_, j_) for i_ in range(N + 1) for j_ in range(N + 1) if i_ in (0, N) or j_ in (0, N)]
                I = np.array([q[0] for q in ij])
                J = np.array([q[1] for q in ij])
                K, Q = np.meshgrid(kk, np.arange(I.size), indexing="ij")
                i, j, K = I[Q.ravel()], J[Q.ravel()], K.ravel()
                a0, a1 = P(i, j, K - 1), P(i, j, K)
                rows.append(dict(p=np.full_like(i, s), f=np.full_like(i, 2), idx=fci(i, j, K), x=0.5 * (a0 + a1),
                                 t=a1 - a0))
    E = {k: np.concatenate([q[k] for q in rows]) for k in rows[0]}
    tree = cKDTree(E["x"])
    parent = np.arange(E["p"].size)

    def root(u):
        while parent[u] != u:
            parent[u] = parent[parent[u]]
            u = parent[u]
        return u
    for u, v in tree.query_pairs(tol):
        ru, rv = root(u), root(v)
        if ru != rv:
            parent[max(ru, rv)] = min(ru, rv)
    roots = np.array([root(u) for u in range(parent.size)])
    order = np.argsort(roots, kind="stable")
    _, starts, counts = np.unique(roots[order], return_index=True, return_counts=True)
    if counts.min() < 2 or counts.max() > 3:
        raise CsError("p1_shell_pairings: the patches' boundary edges do not group by two or three")
    if E["idx"].min() < 0 or E["idx"].max() >= NFC:
        raise CsError("p1_shell_pairings: an edge index is outside the patch layouts")
    for q, n in (("i", NFC), ("z", nz * nz * nzr), ("a", NXF)):
        for side in "ab":
            v = faces[q + side]
            if v.min() < 0 or v.max() >= n:
                raise CsError("p1_shell_pairings: a face index is outside the patch layouts")
    edges = np.zeros(starts.size, EDGE_GROUP)
    edges["n"] = counts
    for m in range(3):
        has = counts > m
        mem = order[np.minimum(starts + m, order.size - 1)]
        first = order[starts]
        edges["p"][:, m] = np.where(has, E["p"][mem], 0)
        edges["f"][:, m] = np.where(has, E["f"][mem], 0)
        edges["idx"][:, m] = np.where(has, E["idx"][mem], 0)
        edges["sign"][:, m] = np.where(has, np.where((E["t"][mem] * E["t"][first]).sum(1) < 0, -1, 1), 0)
    return faces, edges


class P1PatchFluxes(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("Fluct_x_Pl", "Fluct_y_Pl", "Fluct_z_Pl", "Fluct_x_Mi", "Fluct_y_Mi",
                                                "Fluct_z_Mi", "EntFlux_x", "EntFlux_y", "EntFlux_z", "Efx", "Efy",
                                                "Efz", "dU_dt", "xf_area", "yf_area", "zf_area", "v_volume")]


class P1ShellSync(ctypes.Structure):
    _fields_ = [("ntp", ctypes.c_int32), ("nr", ctypes.c_int32), ("nghost", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("patch", P1PatchFluxes * 6), ("faces", ctypes.c_void_p), ("nface", ctypes.c_int64),
                ("edges", ctypes.c_void_p), ("nedge", ctypes.c_int64)]


def sync_shared_p1(N, NR, g, patches, faces_dev, nface, edges_dev, nedge, stream=None):
    """cs_sync_shared_p1 over six patches (dicts of CUDA tensors) with the
    pairings of p1_shell_pairings already on the device."""
    s = P1ShellSync()
    s.ntp, s.nr, s.nghost = N, NR, g
    for q, p in enumerate(patches):
        for k, _ in P1PatchFluxes._fields_:
            setattr(s.patch[q], k, p[k].data_ptr())
    s.faces, s.nface = faces_dev.data_ptr(), nface
    s.edges, s.nedge = edges_dev.data_ptr(), nedge
    st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
    _check(lib().cs_sync_shared_p1(ctypes.byref(s), st), "cs_sync_shared_p1")


# ---------------------------------------------------------------------------
# The reference's time step on the six P1 patches of a shell

class P1Shell:
    """The reference driver's per-step loop over the six P1 patches of one
    shell (1000riemann_geomesh_cs_mhd.f:1746-3461), on the GPU: before each
    stage the ghost exchange (cs_halo_exchange_p1), then per patch
    LIMIT_ALL_SLOPES_P1, MAKE_FACIAL_FLUXES, MAKE_EDGE_EFIELDS_P1,
    BUILD_NON_STIFF_SOURCES; the defect correction in its before-the-update
    form (cs_sync_shared_p1, with defect=True); then per patch
    UPDATE_IMEX_RK_STEP_P1, CLEANUP_AFTER_UPDATE, with the driver's Heun RK2
    bookkeeping (DRV:2645-2773): stage 1 saves U,
    entropy and face B (mode 1) and moves by dt; stage 2 averages them with
    the saved state after the fluxes are formed and moves by dt / 2.
    EVALUATE_TIMESTEP over the patches gives the next step.  With pad=True
    (the driver's single radial patch, DRV:2238-2275) PAD_BOUNDARY_R with
    outflow at both radial ends follows each exchange.

    patches: six dicts of contiguous CUDA tensors holding every array of the
    P1 entry points (cs_p1_patch, cs_p1_flux, cs_p1_edge, cs_p1_source,
    cs_p1_update); scalars: gamma, molewt, rhofloor, prsfloor, smallnum,
    flatten_coef, prs_flatten_coef; sources: (newtonsc, mstar, omega).  The
    patches' face-flux speeds and fluctuations, edge fields, Source and
    dU_dt arrays are work arrays of the step."""

    def __init__(self, cfg, patches, scalars, sources=(1.0, 0.0, (0.0, 0.0, 0.0)), itri=(1, 1, 1, 1),
                 defect=True, pad=True):
        import torch
        if len(patches) != 6:
            raise CsError("P1Shell: six patches (one per cube sector) required")
        if pad and cfg.radial_bc != RBC_PAD:
            # the outflow pads would overwrite U and face B while the entropy
            # kept the wall exchange's mirrored ghosts
            raise CsError("P1Shell: pad=True needs cfg.radial_bc == RBC_PAD (outflow radial pads)")
        self.cfg, self.P, self.sc, self.src, self.itri = cfg, patches, dict(scalars), sources, itri
        self.N, self.NR, self.g = cfg.ntp, cfg.nr, cfg.nghost
        self.solver = CubeSphere(cfg)
        self.saved = [{k: torch.empty_like(p[k][: n]) for k, n in self._mode1(p).items()} for p in patches]
        self.dt_dev = torch.empty(6, dtype=torch.float64, device=patches[0]["U"].device)
        self.defect = defect
        nz = self.N + 2 * self.g
        self.bcarr = torch.full((2 * nz * nz,), 2, dtype=torch.int32, device=patches[0]["U"].device) if pad else None
        if defect:
            geom = [{k: p[k].cpu().numpy() for k in ("vert", "xf_centroid", "yf_centroid", "xf_normal", "yf_normal")}
                    for p in patches]
            faces, edges = p1_shell_pairings(geom, self.N, self.NR, self.g)
            dev = patches[0]["U"].device
            self.faces = torch.from_numpy(faces.view(np.uint8).copy()).to(dev)
            self.edges = torch.from_numpy(edges.view(np.uint8).copy()).to(dev)
            self.nface, self.nedge = faces.size, edges.size

    def _mode1(self, p):
        nz = p["ent"].numel() // 4
        return {"U": 8 * nz, "ent": nz, "Bxf": p["Bxf"].numel(), "Byf": p["Byf"].numel(), "Bzf": p["Bzf"].numel()}

    def exchange(self):
        self.solver.halo_exchange_p1(self.P)
        if self.bcarr is not None:
            for p in self.P:
                pad_boundary_r_p1(self.N, self.NR, self.g, dict(p, bcarr_r=self.bcarr))

    def stage(self, irk, dt):
        N, NR, g = self.N, self.NR, self.g
        self.exchange()
        for p in self.P:
            limit_all_slopes_p1(N, NR, g, self.sc, p, self.itri, minmax=False)
            make_facial_fluxes_p1(N, NR, g, self.sc, p)
            make_edge_efields_p1(N, NR, g, p, self.itri)
            build_non_stiff_sources_p1(N, NR, g, self.src[0], self.src[1], self.sc["smallnum"], self.src[2], p)
        if self.defect:
            sync_shared_p1(N, NR, g, self.P, self.faces, self.nface, self.edges, self.nedge)
        for p, sv in zip(self.P, self.saved):
            if irk == 1:
                for k, n in self._mode1(p).items():
                    sv[k].copy_(p[k][:n])
                move = dt
            else:
                for k, n in self._mode1(p).items():
                    p[k][:n].add_(sv[k]).mul_(0.5)
                move = 0.5 * dt
            update_imex_rk_step_p1(N, NR, g, dict(self.sc, dtmove=move), p)
            cleanup_after_update_p1(N, NR, g, self.sc, p, self.itri)

    def step(self, dt):
        self.stage(1, dt)
        self.stage(2, dt)

    def timestep(self, cfl, dt_old):
        """EVALUATE_TIMESTEP on each patch, the minimum over the six (the
        driver's MPI_ALLREDUCE(MIN), DRV:3515)."""
        for s, p in enumerate(self.P):
            evaluate_timestep_p1(self.N, self.NR, self.g, self.sc, p, cfl, dt_old, out=self.dt_dev[s:s + 1])
        return float(self.dt_dev.min().item())

    def close(self):
        self.solver.close()
