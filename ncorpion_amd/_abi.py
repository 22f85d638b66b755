"""ctypes binding of libncg.so (the C ABI declared in include/ncg.h).

The library is the product; this module only marshals numpy arrays and torch
device tensors into it.  Loading fails loudly when libncg.so has not been
built: there is no fallback implementation.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libncg.so")

NCG_OK, NCG_ERR_INVALID, NCG_ERR_DEVICE, NCG_ERR_NOMEM, NCG_ERR_UNSUPPORTED, NCG_ERR_CAPACITY = range(6)
PRECISIONS = {"exact": 0, "fast64": 1, "mixed": 2}


class Body(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("x", "y", "z", "vx", "vy", "vz", "mass", "radius")]


class Central(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double), ("z", ctypes.c_double),
                ("mass", ctypes.c_double), ("present", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class Pair(ctypes.Structure):
    _fields_ = [("fst", ctypes.c_int32), ("snd", ctypes.c_int32)]


TEST_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32)
RESOLVE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32)
CREATED_FN = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                              ctypes.c_int32)


class ReplayOps(ctypes.Structure):
    _fields_ = [("user", ctypes.c_void_p), ("test", TEST_FN), ("resolve", RESOLVE_FN),
                ("created", CREATED_FN)]


class MeshLoopParams(ctypes.Structure):
    _fields_ = [("mesh_size", ctypes.c_double), ("mesh_size_min", ctypes.c_double),
                ("target_neighbours", ctypes.c_double), ("r_disruption", ctypes.c_double),
                ("r_high_dumping", ctypes.c_double), ("cells", ctypes.c_int32),
                ("host_passes", ctypes.c_int32), ("conflict_passes", ctypes.c_int32)]


class BuiltinCollider(ctypes.Structure):
    _fields_ = [("bodies", ctypes.c_void_p), ("exists", ctypes.c_void_p),
                ("did_collide", ctypes.c_void_p), ("timestep", ctypes.c_double),
                ("collision_f", ctypes.c_double), ("one_collision_only", ctypes.c_int32),
                ("pad_", ctypes.c_int32), ("collisions", ctypes.c_int64),
                ("approach_", ctypes.c_double * 6), ("t_c_", ctypes.c_double)]


class BuiltinOps:
    """Library-side closest_approach + bounce on a host body array (kept
    alive with the arrays it points to)."""

    def __init__(self, lib, bodies, exists, did_collide, dt, f, one_collision_only):
        self.bodies, self.exists, self.did_collide = bodies, exists, did_collide
        self.state = BuiltinCollider()
        self.state.bodies = bodies.ctypes.data
        self.state.exists = None if exists is None else exists.ctypes.data
        self.state.did_collide = None if did_collide is None else did_collide.ctypes.data
        self.state.timestep = float(dt)
        self.state.collision_f = float(f)
        self.state.one_collision_only = int(one_collision_only)
        self.ops = ReplayOps()
        st = lib.ncg_builtin_replay_ops(ctypes.byref(self.state), ctypes.byref(self.ops))
        if st != NCG_OK:
            raise NcgError(st, "ncg_builtin_replay_ops failed")

    @property
    def collisions(self):
        return int(self.state.collisions)


OUTCOMES = {"bounce": 0, "merger": 1, "fragmentation": 2}
MAX_TILDE = 63


class OutcomeModel(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("n_tilde", ctypes.c_int32), ("pq", ctypes.c_int32 * 4),
                ("collision_f", ctypes.c_double), ("mu", ctypes.c_double), ("nu", ctypes.c_double),
                ("c1", ctypes.c_double), ("k", ctypes.c_double),
                ("merge_threshold", ctypes.c_double), ("frag_threshold", ctypes.c_double),
                ("grav_const", ctypes.c_double)]


def outcome_model(kind, collision_f=1.5, G=1.0, n_tilde=15, pq=(-1, 3, -1, 1), mu=0.55, nu=0.4,
                  c1=1.5, k=0.2, merging_threshold=0.0002, fragment_threshold=1.5e-8):
    """An OutcomeModel; the defaults are NcorpiON's parameters.h values."""
    m = OutcomeModel()
    m.kind = OUTCOMES[kind]
    m.n_tilde = int(n_tilde)
    for q in range(4):
        m.pq[q] = int(pq[q])
    m.collision_f = float(collision_f)
    m.mu, m.nu, m.c1, m.k = float(mu), float(nu), float(c1), float(k)
    m.merge_threshold = float(merging_threshold)
    m.frag_threshold = float(fragment_threshold)
    m.grav_const = float(G)
    return m


class LeapfrogModelParams(ctypes.Structure):
    _fields_ = [("model", OutcomeModel), ("r_disruption", ctypes.c_double),
                ("r_high_dumping", ctypes.c_double), ("reduce_to_com", ctypes.c_int32),
                ("n_max", ctypes.c_int32), ("largest_id", ctypes.c_int32),
                ("how_many_free", ctypes.c_int32), ("free_indexes", ctypes.c_void_p),
                ("need_to_reduce_com", ctypes.c_int32), ("host_passes", ctypes.c_int32),
                ("host_step_ends", ctypes.c_int32), ("removed", ctypes.c_int32),
                ("tidy_ups", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("mergers", ctypes.c_int64), ("super_catastrophic", ctypes.c_int64),
                ("partial", ctypes.c_int64), ("full", ctypes.c_int64)]


class MeshStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("total_neighbours", "how_many_small", "how_many_big",
                                             "collisions", "candidates", "host_tests",
                                             "replay_conflicts")]


_lib = None

P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
D = ctypes.c_double
S = ctypes.c_int

_SIGNATURES = {
    "ncg_abi_version": (S, []),
    "ncg_create": (S, [ctypes.POINTER(P), S]),
    "ncg_destroy": (None, [P]),
    "ncg_last_error": (ctypes.c_char_p, [P]),
    "ncg_synchronize": (S, [P]),
    "ncg_stream": (P, [P]),
    "ncg_gravity_allpairs": (S, [P, P, I32, P, ctypes.POINTER(Central), D, D, S, P, P]),
    "ncg_gravity_allpairs_dev": (S, [P, P, I32, P, ctypes.POINTER(Central), D, D, S, P, P, P]),
    "ncg_gravity_allpairs_static_dev": (S, [P, P, I32, P, ctypes.POINTER(Central), D, D, S, P, P,
                                            P]),
    "ncg_brute_force_pass": (S, [P, P, I32, P, D, ctypes.POINTER(ReplayOps), ctypes.POINTER(I64)]),
    "ncg_collide_candidates_allpairs": (S, [P, P, I32, P, D, P, I64, ctypes.POINTER(I64)]),
    "ncg_mesh_pass": (S, [P, P, I32, P, D, D, D, I32, I32, ctypes.POINTER(ReplayOps),
                          ctypes.POINTER(MeshStats)]),
    "ncg_mesh_gravity": (S, [P, P, I32, P, ctypes.POINTER(Central), D, D, S, P, P]),
    "ncg_mesh_gravity_dev": (S, [P, P, I32, P, P, D, D, S, P, P, P]),
    "ncg_gravity_allpairs_from": (S, [P, P, I32, P, P, D, D, S, P]),
    "ncg_mesh_gravity_from": (S, [P, P, I32, P, P, D, D, S, P]),
    "ncg_mesh_leapfrog_dev": (S, [P, P, I32, P, P, P, D, D, S, D, I32,
                                  ctypes.POINTER(MeshLoopParams), D, I32, I32,
                                  ctypes.POINTER(I64), P]),
    "ncg_mesh_pass_dev": (S, [P, P, I32, P, P, D, D, D, I32, I32, D, I32,
                              ctypes.POINTER(MeshStats), P]),
    "ncg_mesh_pairs": (S, [P, P, I64, ctypes.POINTER(I64)]),
    "ncg_rk4_steps_dev": (S, [P, P, I32, P, P, D, D, S, D, I32, I32, P]),
    "ncg_set_tiling": (S, [P, I32, I32, D, I32]),
    "ncg_vjp_gravity_dev": (S, [P, P, I32, P, D, D, P, P, P, P, P, P]),
    "ncg_rk4_adjoint_dev": (S, [P, P, I32, P, D, D, D, I32, I32, P, P, P, P, P, P, P]),
    "ncg_ensemble_rk4_dev": (S, [P, P, P, I32, P, D, D, S, D, I32, I32, D, ctypes.POINTER(I64), P]),
    "ncg_ensemble_gravity_dev": (S, [P, P, P, I32, P, D, D, S, P, P, P]),
    "ncg_energy_dev": (S, [P, P, I32, P, P, D, D, P, P]),
    "ncg_leapfrog_dev": (S, [P, P, I32, P, P, P, D, D, S, D, I32, D, I32, I32,
                             ctypes.POINTER(I64), P]),
    "ncg_leapfrog_model_dev": (S, [P, P, P, P, P, D, D, S, D, I32, I32, I32,
                                   ctypes.POINTER(LeapfrogModelParams), ctypes.POINTER(I64), P]),
    "ncg_mesh_leapfrog_model_dev": (S, [P, P, P, P, P, D, D, S, D, I32,
                                        ctypes.POINTER(MeshLoopParams),
                                        ctypes.POINTER(LeapfrogModelParams), I32, I32,
                                        ctypes.POINTER(I64), P]),
    "ncg_rk4_forward_ckpt_dev": (S, [P, P, I32, P, D, D, D, I32, I32, P, P, P]),
    "ncg_rk4_adjoint_ckpt_dev": (S, [P, P, I32, P, D, D, D, I32, I32, P, P, P, P, P, P]),
    "ncg_builtin_replay_ops": (S, [ctypes.POINTER(BuiltinCollider), ctypes.POINTER(ReplayOps)]),
    "ncg_builtin_outcome_ops": (S, [P, ctypes.POINTER(ReplayOps)]),
    "ncg_format_fixed": (S, [P, P, I64, I32, P, I64, ctypes.POINTER(I64)]),
    "ncg_cart2ell": (S, [P, P, I32, P, P, D, D, P, P]),
    "ncg_profile_enable": (S, [P, S]),
    "ncg_profile_read": (S, [P, ctypes.c_char_p, ctypes.POINTER(D), ctypes.POINTER(I64), S]),
}


def load():
    """Load libncg.so and declare the C signatures (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libncg.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    # PyTorch-ROCm ships its own HIP runtime with the same SONAME
    # (libamdhip64.so.7).  Loading torch first makes libncg bind to that one,
    # so torch tensors and libncg share a single runtime in this process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class NcgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("ncg status %d: %s" % (status, msg))
        self.status = status


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def as_bodies(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != 8:
        raise ValueError("bodies must have shape (n, 8)")
    return a


def as_exists(e, n):
    if e is None:
        return None
    e = np.ascontiguousarray(e, dtype=np.int32)
    if e.shape != (n,):
        raise ValueError("exists must have shape (n,)")
    return e


def make_central(central):
    """central: None, or (x, y, z, M)."""
    c = Central()
    if central is not None:
        c.x, c.y, c.z, c.mass = [float(v) for v in central]
        c.present = 1
    return c


class Context:
    """One GPU context of libncg (ncg_create / ncg_destroy)."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        st = self.lib.ncg_create(ctypes.byref(h), int(device))
        if st != NCG_OK:
            raise NcgError(st, "ncg_create(device=%d) failed" % device)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.ncg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != NCG_OK:
            raise NcgError(st, self.lib.ncg_last_error(self.h).decode(errors="replace"))

    @property
    def stream(self):
        return self.lib.ncg_stream(self.h)

    def synchronize(self):
        self._check(self.lib.ncg_synchronize(self.h))

    # -- brute-force gravity (vector_field) ------------------------------------
    def gravity_allpairs(self, bodies, exists=None, central=None, G=1.0, softening=0.0,
                         precision="exact", acc=None):
        b = as_bodies(bodies)
        n = b.shape[0]
        e = as_exists(exists, n)
        out = np.zeros((n, 3)) if acc is None else acc
        cm = np.zeros(3)
        c = make_central(central)
        self._check(self.lib.ncg_gravity_allpairs(self.h, _ptr(b), n, _ptr(e), ctypes.byref(c),
                                                  float(G), float(softening),
                                                  PRECISIONS[precision], _ptr(out), _ptr(cm)))
        return out, cm

    def gravity_allpairs_from(self, bodies, start, exists=None, G=1.0, softening=0.0,
                              precision="exact"):
        """Mutual accelerations added to the caller's start values (3 per body;
        see ncg_gravity_allpairs_from)."""
        b = as_bodies(bodies)
        n = b.shape[0]
        e = as_exists(exists, n)
        st = np.ascontiguousarray(start, dtype=np.float64).reshape(n, 3)
        out = np.zeros((n, 3))
        self._check(self.lib.ncg_gravity_allpairs_from(self.h, _ptr(b), n, _ptr(e), _ptr(st),
                                                       float(G), float(softening),
                                                       PRECISIONS[precision], _ptr(out)))
        return out

    def gravity_allpairs_dev(self, bodies_t, acc_t, exists_t=None, central=None, G=1.0,
                             softening=0.0, precision="mixed", cm_t=None, stream=None):
        """bodies_t: (n, 8) float64 device tensor; acc_t: (n, 3) float64 device tensor."""
        n = int(bodies_t.shape[0])
        c = make_central(central)
        self._check(self.lib.ncg_gravity_allpairs_dev(
            self.h, bodies_t.data_ptr(), n, None if exists_t is None else exists_t.data_ptr(),
            ctypes.byref(c), float(G), float(softening), PRECISIONS[precision], acc_t.data_ptr(),
            None if cm_t is None else cm_t.data_ptr(), stream))

    def gravity_allpairs_static_dev(self, bodies_t, acc_t, exists_t=None, central=None, G=1.0,
                                    softening=0.0, precision="mixed", cm_t=None, stream=None):
        """gravity_allpairs_dev through the allocation-free entry point (its
        scratch is the library's static per-GPU array; n <= 32768)."""
        n = int(bodies_t.shape[0])
        c = make_central(central)
        self._check(self.lib.ncg_gravity_allpairs_static_dev(
            self.h, bodies_t.data_ptr(), n, None if exists_t is None else exists_t.data_ptr(),
            ctypes.byref(c), float(G), float(softening), PRECISIONS[precision], acc_t.data_ptr(),
            None if cm_t is None else cm_t.data_ptr(), stream))

    def rk4_steps_dev(self, bodies_t, central_t, nsteps, dt, G=1.0, softening=0.0,
                      precision="mixed", exists_t=None, graph=True, stream=None):
        """Advance bodies_t ((n, 8) float64, device) and central_t (8 float64,
        device: x, y, z, vx, vy, vz, M, present) by nsteps RK4 steps."""
        n = int(bodies_t.shape[0])
        self._check(self.lib.ncg_rk4_steps_dev(
            self.h, bodies_t.data_ptr(), n, None if exists_t is None else exists_t.data_ptr(),
            central_t.data_ptr(), float(G), float(softening), PRECISIONS[precision], float(dt),
            int(nsteps), 1 if graph else 0, stream))

    def ensemble_rk4_dev(self, bodies_t, offsets, centrals_t, nsteps, dt, G=1.0, softening=0.0,
                         precision="mixed", graph=True, collisions=False, collision_f=1.5,
                         stream=None):
        """bodies_t: (sum N, 8) float64 device; offsets: host int sequence of
        nsys+1; centrals_t: (nsys, 8) float64 device.  Returns collisions."""
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        nc = I64()
        flags = (1 if graph else 0) | (2 if collisions else 0)
        self._check(self.lib.ncg_ensemble_rk4_dev(
            self.h, bodies_t.data_ptr(), _ptr(off), len(off) - 1, centrals_t.data_ptr(),
            float(G), float(softening), PRECISIONS[precision], float(dt), int(nsteps), flags,
            float(collision_f), ctypes.byref(nc), stream))
        return int(nc.value)

    def ensemble_gravity_dev(self, bodies_t, offsets, centrals_t, acc_t, cm_t, G=1.0,
                             softening=0.0, precision="mixed", stream=None):
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        self._check(self.lib.ncg_ensemble_gravity_dev(
            self.h, bodies_t.data_ptr(), _ptr(off), len(off) - 1, centrals_t.data_ptr(),
            float(G), float(softening), PRECISIONS[precision], acc_t.data_ptr(), cm_t.data_ptr(),
            stream))

    def vjp_gravity_dev(self, bodies_t, central_t, lam_t, gx_t, gm_t, gc_t, lam_c_t=None, G=1.0,
                        softening=0.0, stream=None):
        n = int(bodies_t.shape[0])
        self._check(self.lib.ncg_vjp_gravity_dev(
            self.h, bodies_t.data_ptr(), n, central_t.data_ptr(), float(G), float(softening),
            lam_t.data_ptr(), None if lam_c_t is None else lam_c_t.data_ptr(), gx_t.data_ptr(),
            gm_t.data_ptr(), gc_t.data_ptr(), stream))

    def leapfrog_dev(self, bodies_t, central_t, nsteps, dt, G=1.0, softening=0.0,
                     precision="exact", collision_f=0.0, one_collision_only=0, exists_t=None,
                     did_collide_t=None, graph=False, half_drift_start=False,
                     half_drift_end=False, stream=None):
        """NcorpiON's brute-force leapfrog on the device (see ncg_leapfrog_dev);
        returns the number of collisions resolved."""
        n = int(bodies_t.shape[0])
        nc = I64()
        flags = (1 if graph else 0) | (2 if half_drift_start else 0) | (4 if half_drift_end else 0)
        self._check(self.lib.ncg_leapfrog_dev(
            self.h, bodies_t.data_ptr(), n, None if exists_t is None else exists_t.data_ptr(),
            None if did_collide_t is None else did_collide_t.data_ptr(), central_t.data_ptr(),
            float(G), float(softening), PRECISIONS[precision], float(dt), int(nsteps),
            float(collision_f), int(one_collision_only), flags, ctypes.byref(nc), stream))
        return int(nc.value)

    def leapfrog_model_dev(self, bodies_t, central_t, exists_t, nsteps, dt, state, model,
                           G=1.0, softening=0.0, precision="exact", one_collision_only=0,
                           did_collide_t=None, r_disruption=0.0, r_high_dumping=0.0,
                           reduce_to_com=1, graph=False, half_drift_start=False,
                           half_drift_end=False, stream=None):
        """NcorpiON's brute-force leapfrog with an outcome model and
        end_of_timestep's removals (see ncg_leapfrog_model_dev).  bodies_t,
        exists_t, did_collide_t: device tensors of n_max slots.  state: dict
        with largest_id, how_many_free, free_indexes (int32 numpy array of
        n_max) and need_to_reduce_com, updated in place; the call's counters
        are added under their names.  Returns the collisions resolved."""
        n_max = int(bodies_t.shape[0])
        mp = self._model_params(n_max, state, model, r_disruption, r_high_dumping, reduce_to_com)
        nc = I64()
        flags = (1 if graph else 0) | (2 if half_drift_start else 0) | (4 if half_drift_end else 0)
        self._check(self.lib.ncg_leapfrog_model_dev(
            self.h, bodies_t.data_ptr(), exists_t.data_ptr(),
            None if did_collide_t is None else did_collide_t.data_ptr(), central_t.data_ptr(),
            float(G), float(softening), PRECISIONS[precision], float(dt), int(nsteps),
            int(one_collision_only), flags, ctypes.byref(mp), ctypes.byref(nc), stream))
        self._model_state_back(mp, state)
        return int(nc.value)

    @staticmethod
    def _model_params(n_max, state, model, r_disruption, r_high_dumping, reduce_to_com):
        free = state["free_indexes"]
        assert free.dtype == np.int32 and free.shape[0] == n_max and free.flags["C_CONTIGUOUS"]
        mp = LeapfrogModelParams()
        mp.model = model
        mp.r_disruption = float(r_disruption)
        mp.r_high_dumping = float(r_high_dumping)
        mp.reduce_to_com = int(reduce_to_com)
        mp.n_max = n_max
        mp.largest_id = int(state["largest_id"])
        mp.how_many_free = int(state["how_many_free"])
        mp.free_indexes = free.ctypes.data
        mp.need_to_reduce_com = int(state.get("need_to_reduce_com", 0))
        return mp

    @staticmethod
    def _model_state_back(mp, state):
        state["largest_id"] = int(mp.largest_id)
        state["how_many_free"] = int(mp.how_many_free)
        state["need_to_reduce_com"] = int(mp.need_to_reduce_com)
        for f in ("host_passes", "host_step_ends", "removed", "tidy_ups", "mergers",
                  "super_catastrophic", "partial", "full"):
            state[f] = state.get(f, 0) + int(getattr(mp, f))

    def mesh_leapfrog_model_dev(self, bodies_t, central_t, exists_t, nsteps, dt, params, state,
                                model, G=1.0, softening=0.0, precision="exact",
                                one_collision_only=0, did_collide_t=None, r_disruption=0.0,
                                r_high_dumping=0.0, reduce_to_com=1, half_kick_start=False,
                                half_kick_end=False, stream=None):
        """NcorpiON's mesh integrator with an outcome model and end_of_timestep's
        removals (see ncg_mesh_leapfrog_model_dev).  params: as
        mesh_leapfrog_dev (gam updated in place); state, model: as
        leapfrog_model_dev.  Returns the collisions resolved."""
        n_max = int(bodies_t.shape[0])
        mesh = MeshLoopParams(float(params["gam"]), float(params["gam_min"]),
                              float(params["how_many_neighbours"]), 0.0, 0.0,
                              int(params["cells"]), 0, 0)
        mp = self._model_params(n_max, state, model, r_disruption, r_high_dumping, reduce_to_com)
        nc = I64()
        flags = (2 if half_kick_start else 0) | (4 if half_kick_end else 0)
        self._check(self.lib.ncg_mesh_leapfrog_model_dev(
            self.h, bodies_t.data_ptr(), exists_t.data_ptr(),
            None if did_collide_t is None else did_collide_t.data_ptr(), central_t.data_ptr(),
            float(G), float(softening), PRECISIONS[precision], float(dt), int(nsteps),
            ctypes.byref(mesh), ctypes.byref(mp), int(one_collision_only), flags,
            ctypes.byref(nc), stream))
        params["gam"] = mesh.mesh_size
        params["host_passes"] = int(mesh.host_passes)
        params["conflict_passes"] = int(mesh.conflict_passes)
        self._model_state_back(mp, state)
        return int(nc.value)

    def mesh_pass_dev(self, bodies_t, dt, gam, cube, cells, mode, collision_f=1.5,
                      one_collision_only=0, exists_t=None, did_collide_t=None, stream=None):
        """mesh() (mode 0, device replay with the library's bounce) or
        get_neighbours_mesh() (mode 1) on device tensors; returns the counters."""
        n = int(bodies_t.shape[0])
        stats = MeshStats()
        self._check(self.lib.ncg_mesh_pass_dev(
            self.h, bodies_t.data_ptr(), n, None if exists_t is None else exists_t.data_ptr(),
            None if did_collide_t is None else did_collide_t.data_ptr(), float(dt), float(gam),
            float(cube), int(cells), int(mode), float(collision_f), int(one_collision_only),
            ctypes.byref(stats), stream))
        return {f: int(getattr(stats, f)) for f, _ in MeshStats._fields_}

    def mesh_leapfrog_dev(self, bodies_t, central_t, nsteps, dt, params, G=1.0, softening=0.0,
                          precision="exact", collision_f=0.0, one_collision_only=0,
                          exists_t=None, did_collide_t=None, half_kick_start=False,
                          half_kick_end=False, stream=None):
        """NcorpiON's mesh integrator on the device (see ncg_mesh_leapfrog_dev).
        params: dict with gam, gam_min, how_many_neighbours, disruption_threshold,
        high_dumping_threshold, cells; params["gam"] is updated in place.
        Returns the number of collisions resolved."""
        n = int(bodies_t.shape[0])
        mp = MeshLoopParams(float(params["gam"]), float(params["gam_min"]),
                            float(params["how_many_neighbours"]),
                            float(params["disruption_threshold"]),
                            float(params["high_dumping_threshold"]), int(params["cells"]), 0, 0)
        nc = I64()
        flags = (2 if half_kick_start else 0) | (4 if half_kick_end else 0)
        self._check(self.lib.ncg_mesh_leapfrog_dev(
            self.h, bodies_t.data_ptr(), n, None if exists_t is None else exists_t.data_ptr(),
            None if did_collide_t is None else did_collide_t.data_ptr(), central_t.data_ptr(),
            float(G), float(softening), PRECISIONS[precision], float(dt), int(nsteps),
            ctypes.byref(mp), float(collision_f), int(one_collision_only), flags,
            ctypes.byref(nc), stream))
        params["gam"] = mp.mesh_size
        params["host_passes"] = int(mp.host_passes)
        params["conflict_passes"] = int(mp.conflict_passes)
        return int(nc.value)

    def mesh_gravity_dev(self, bodies_t, acc_t, central_t=None, cm_t=None, exists_t=None, G=1.0,
                         softening=0.0, precision="exact", stream=None):
        """Mesh field of device tensors over the last mesh pass (asynchronous)."""
        n = int(bodies_t.shape[0])
        self._check(self.lib.ncg_mesh_gravity_dev(
            self.h, bodies_t.data_ptr(), n, None if exists_t is None else exists_t.data_ptr(),
            None if central_t is None else central_t.data_ptr(), float(G), float(softening),
            PRECISIONS[precision], acc_t.data_ptr(), None if cm_t is None else cm_t.data_ptr(),
            stream))

    def energy_dev(self, bodies_t, central_t=None, exists_t=None, G=1.0, softening=0.0,
                   out_t=None, stream=None):
        """Energy terms {kinetic, mutual potential, central potential, total}
        as a (4,) float64 device tensor (written asynchronously on `stream`)."""
        import torch
        n = int(bodies_t.shape[0])
        if out_t is None:
            out_t = torch.empty(4, dtype=torch.float64, device=bodies_t.device)
        self._check(self.lib.ncg_energy_dev(
            self.h, bodies_t.data_ptr(), n, None if exists_t is None else exists_t.data_ptr(),
            None if central_t is None else central_t.data_ptr(), float(G), float(softening),
            out_t.data_ptr(), stream))
        return out_t

    def rk4_forward_ckpt_dev(self, bodies_t, central_t, nsteps, dt, checkpoint_every=20, G=1.0,
                             softening=0.0, ckpt_t=None, ckpt_central_t=None, stream=None):
        """Advances bodies_t / central_t in place by nsteps FP64 RK4 steps and
        returns (ckpt_t, ckpt_central_t), the states every checkpoint_every
        steps (allocated when not given)."""
        import torch
        n = int(bodies_t.shape[0])
        nck = max(1, (int(nsteps) - 1) // int(checkpoint_every) + 1)
        if ckpt_t is None:
            ckpt_t = torch.empty((nck, n, 8), dtype=torch.float64, device=bodies_t.device)
        if ckpt_central_t is None:
            ckpt_central_t = torch.empty((nck, 8), dtype=torch.float64, device=bodies_t.device)
        self._check(self.lib.ncg_rk4_forward_ckpt_dev(
            self.h, bodies_t.data_ptr(), n, central_t.data_ptr(), float(G), float(softening),
            float(dt), int(nsteps), int(checkpoint_every), ckpt_t.data_ptr(),
            ckpt_central_t.data_ptr(), stream))
        return ckpt_t, ckpt_central_t

    def rk4_adjoint_ckpt_dev(self, ckpt_t, ckpt_central_t, gfinal_t, g0_t, gm_t, gc0_t, nsteps, dt,
                             checkpoint_every=20, G=1.0, softening=0.0, gcentral_final_t=None,
                             stream=None):
        n = int(ckpt_t.shape[1])
        self._check(self.lib.ncg_rk4_adjoint_ckpt_dev(
            self.h, ckpt_t.data_ptr(), n, ckpt_central_t.data_ptr(), float(G), float(softening),
            float(dt), int(nsteps), int(checkpoint_every), gfinal_t.data_ptr(),
            None if gcentral_final_t is None else gcentral_final_t.data_ptr(), g0_t.data_ptr(),
            gm_t.data_ptr(), gc0_t.data_ptr(), stream))

    def rk4_adjoint_dev(self, bodies0_t, central0_t, gfinal_t, g0_t, gm_t, gc0_t, nsteps, dt,
                        checkpoint_every=20, G=1.0, softening=0.0, gcentral_final_t=None,
                        final_t=None, stream=None):
        n = int(bodies0_t.shape[0])
        self._check(self.lib.ncg_rk4_adjoint_dev(
            self.h, bodies0_t.data_ptr(), n, central0_t.data_ptr(), float(G), float(softening),
            float(dt), int(nsteps), int(checkpoint_every), gfinal_t.data_ptr(),
            None if gcentral_final_t is None else gcentral_final_t.data_ptr(), g0_t.data_ptr(),
            gm_t.data_ptr(), gc0_t.data_ptr(), None if final_t is None else final_t.data_ptr(),
            stream))

    def set_tiling(self, ipl=0, nsplit=-1, theta=0.0, source=-1):
        self._check(self.lib.ncg_set_tiling(self.h, int(ipl), int(nsplit), float(theta),
                                            int(source)))

    def profile(self, enable=True):
        self._check(self.lib.ncg_profile_enable(self.h, 1 if enable else 0))

    def profile_read(self, kernel, reset=True):
        ms = D()
        cnt = I64()
        self._check(self.lib.ncg_profile_read(self.h, kernel.encode(), ctypes.byref(ms),
                                              ctypes.byref(cnt), 1 if reset else 0))
        return float(ms.value), int(cnt.value)

    # -- brute-force collisions (brute_force) ----------------------------------
    def collide_candidates(self, bodies, exists=None, dt=0.0):
        b = as_bodies(bodies)
        n = b.shape[0]
        e = as_exists(exists, n)
        cap = 1 << 16
        while True:
            out = np.zeros((cap, 2), dtype=np.int32)
            cnt = I64()
            st = self.lib.ncg_collide_candidates_allpairs(self.h, _ptr(b), n, _ptr(e), float(dt),
                                                          _ptr(out), cap, ctypes.byref(cnt))
            if st == NCG_ERR_CAPACITY:
                cap = int(cnt.value)
                continue
            self._check(st)
            return out[: cnt.value].copy()

    def builtin_ops(self, bodies, dt, exists=None, did_collide=None, f=1.5, one_collision_only=0):
        """Built-in decision/outcome ops (ncg_builtin_replay_ops) acting on the
        host array `bodies` in place; pass as ops= to brute_force_pass or
        mesh_pass."""
        if not (isinstance(bodies, np.ndarray) and bodies.dtype == np.float64
                and bodies.flags.c_contiguous):
            raise ValueError("bodies must be a C-contiguous float64 array (mutated in place)")
        n = bodies.shape[0]
        e = None if exists is None else as_exists(exists, n)
        dc = None
        if did_collide is not None:
            dc = did_collide
            if not (isinstance(dc, np.ndarray) and dc.dtype == np.int32 and dc.flags.c_contiguous):
                raise ValueError("did_collide must be a C-contiguous int32 array")
        return BuiltinOps(self.lib, bodies, e, dc, dt, f, one_collision_only)

    def _ops(self, test, resolve, ops, created=None):
        """(keepalive, pointer) of the replay ops: built-in or Python callbacks.
        created() returns the slots brought into existence by the last resolve."""
        if ops is not None:
            return ops, ctypes.byref(ops.ops)
        if test is None:
            return None, None
        tf = TEST_FN(lambda u, a, c: 1 if test(int(a), int(c)) else 0)
        rf = RESOLVE_FN(lambda u, a, c: resolve(int(a), int(c)))

        def _created(u, ids, cap):
            new = list(created())
            if len(new) > cap:
                return -1
            for q, v in enumerate(new):
                ids[q] = int(v)
            return len(new)

        cf = CREATED_FN(_created) if created is not None else CREATED_FN()
        r = ReplayOps(None, tf, rf, cf)
        return (r, tf, rf, cf), ctypes.byref(r)

    def brute_force_pass(self, bodies, exists, dt, test=None, resolve=None, ops=None,
                         created=None, n=None):
        """Replay of brute_force with Python callbacks test(a, b) -> bool,
        resolve(a, b) and, when outcomes create bodies, created() -> new slots;
        or built-in ops.  `bodies` (and `exists`) must be the arrays they
        mutate; n (default: all rows) is the number of slots of the pass, the
        arrays may be longer to hold bodies created during it."""
        b = bodies
        if not (isinstance(b, np.ndarray) and b.dtype == np.float64 and b.flags.c_contiguous):
            raise ValueError("bodies must be a C-contiguous float64 array (mutated in place)")
        n = b.shape[0] if n is None else int(n)
        if n > b.shape[0]:
            raise ValueError("n exceeds the body array")
        if exists is not None:
            if not (isinstance(exists, np.ndarray) and exists.dtype == np.int32
                    and exists.flags.c_contiguous and exists.shape == (b.shape[0],)):
                raise ValueError("exists must be a C-contiguous int32 array of one entry per "
                                 "body slot (the live array the callbacks mutate)")
        e = exists
        keep, ops_p = self._ops(test, resolve, ops, created)
        if ops_p is None:
            raise ValueError("brute_force_pass needs test/resolve callbacks or ops")
        nc = I64()
        self._check(self.lib.ncg_brute_force_pass(self.h, _ptr(b), n, _ptr(e), float(dt),
                                                  ops_p, ctypes.byref(nc)))
        del keep
        return int(nc.value)

    # -- mesh module -------------------------------------------------------------
    def mesh_pass(self, bodies, exists, dt, gam, cube, cells, mode, test=None, resolve=None,
                  ops=None, created=None, n=None):
        """mesh() (mode 0) or get_neighbours_mesh() (mode 1).  Mode 0 takes
        Python callbacks test(a, b) -> bool, resolve(a, b) and, when outcomes
        append bodies, created() -> new slots; or built-in ops.  n (default:
        all rows) is the number of slots of the pass; the arrays may be longer
        to hold the bodies appended during it, and `exists` must then be the
        live int32 array the callbacks mutate."""
        b = bodies
        if not (isinstance(b, np.ndarray) and b.dtype == np.float64 and b.flags.c_contiguous):
            raise ValueError("bodies must be a C-contiguous float64 array (mutated in place)")
        if n is None:
            n = b.shape[0]
            e = as_exists(exists, n)
        else:
            n = int(n)
            if n > b.shape[0]:
                raise ValueError("n exceeds the body array")
            e = exists
            if e is not None and not (isinstance(e, np.ndarray) and e.dtype == np.int32
                                      and e.flags.c_contiguous and e.shape == (b.shape[0],)):
                raise ValueError("exists must be a C-contiguous int32 array of one entry per "
                                 "body slot (the live array the callbacks mutate)")
        stats = MeshStats()
        keep, ops_p = self._ops(test, resolve, ops, created)
        self._check(self.lib.ncg_mesh_pass(self.h, _ptr(b), n, _ptr(e), float(dt), float(gam),
                                           float(cube), int(cells), int(mode), ops_p,
                                           ctypes.byref(stats)))
        del keep
        return {f: int(getattr(stats, f)) for f, _ in MeshStats._fields_}

    def mesh_gravity(self, bodies, exists=None, central=None, G=1.0, softening=0.0,
                     precision="exact"):
        b = as_bodies(bodies)
        n = b.shape[0]
        e = as_exists(exists, n)
        out = np.zeros((n, 3))
        cm = np.zeros(3)
        c = make_central(central)
        self._check(self.lib.ncg_mesh_gravity(self.h, _ptr(b), n, _ptr(e), ctypes.byref(c),
                                              float(G), float(softening), PRECISIONS[precision],
                                              _ptr(out), _ptr(cm)))
        return out, cm

    def mesh_gravity_from(self, bodies, start, exists=None, G=1.0, softening=0.0,
                          precision="exact"):
        """ncg_mesh_gravity_from: the last pass's field added to start values."""
        b = as_bodies(bodies)
        n = b.shape[0]
        e = as_exists(exists, n)
        st = np.ascontiguousarray(start, dtype=np.float64).reshape(n, 3)
        out = np.zeros((n, 3))
        self._check(self.lib.ncg_mesh_gravity_from(self.h, _ptr(b), n, _ptr(e), _ptr(st),
                                                   float(G), float(softening),
                                                   PRECISIONS[precision], _ptr(out)))
        return out

    # -- output path -------------------------------------------------------------
    def format_fixed(self, values, decimals=14):
        """Text of printf("%.<decimals>f ") for each value, concatenated (bytes)."""
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        cap = max(64, 24 * v.size)
        while True:
            out = ctypes.create_string_buffer(cap)
            ln = I64()
            st = self.lib.ncg_format_fixed(self.h, _ptr(v), v.size, int(decimals), out, cap,
                                           ctypes.byref(ln))
            if st == NCG_ERR_CAPACITY:
                cap = int(ln.value)
                continue
            self._check(st)
            return out.raw[: ln.value]

    def cart2ell(self, bodies, exists=None, central=None, mu=1.0, G=1.0):
        """(alkhqp (n, 6), warn (n,)) of cart2ell for every existing body;
        central: (x, y, z, vx, vy, vz) or None."""
        b = as_bodies(bodies)
        n = b.shape[0]
        e = as_exists(exists, n)
        c = None if central is None else np.ascontiguousarray(central, dtype=np.float64)
        out = np.zeros((n, 6))
        warn = np.zeros(n, np.int32)
        self._check(self.lib.ncg_cart2ell(self.h, _ptr(b), n, _ptr(e), _ptr(c), float(mu), float(G),
                                          _ptr(out), _ptr(warn)))
        return out, warn

    def mesh_pairs(self):
        cap = 1 << 20
        while True:
            out = np.zeros((cap, 2), dtype=np.int32)
            cnt = I64()
            st = self.lib.ncg_mesh_pairs(self.h, _ptr(out), cap, ctypes.byref(cnt))
            if st == NCG_ERR_CAPACITY:
                cap = int(cnt.value)
                continue
            self._check(st)
            return out[: cnt.value].copy()
