"""bonsai2_amd -- MI355X Barnes-Hut treecode (host-side Python mirror).

``Treecode`` mirrors the reference class Treecode<float>
(code/Treecode.h:123-223): the same method names, argument meanings and call
order (alloc -> particles -> buildTree -> computeMultipoles -> makeGroups ->
computeForces -> moveParticles / computeEnergies).  Every call goes through
the C ABI of libbonsai2_amd.so (include/bonsai2_amd.h), which runs the
hand-written HIP kernels; there is no CPU fallback -- if the library is not
built, importing the binding raises.
"""
import ctypes
import os

import numpy as np

__all__ = ["Treecode", "lib", "BonsaiError", "ORDER_INPUT", "ORDER_TREE", "ORDER_PH",
           "plummer", "clustered", "halo_disk", "ensemble", "read_tipsy", "write_tipsy",
           "EXPORTED_SYMBOLS"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbonsai2_amd.so")

ORDER_INPUT, ORDER_TREE, ORDER_PH = 0, 1, 2
# radix passes per sort of the 63-bit keys (bytes per pass: 24 per key-value pair)
SORT_PASSES = 8
P = ctypes.c_void_p
I = ctypes.c_int
F = ctypes.c_float
D = ctypes.c_double

# name -> argtypes (restype int for all)
_SIGS = {
    "bt_create": [ctypes.POINTER(P), F, F, I],
    "bt_destroy": [P],
    "bt_alloc": [P, I, I, P],
    "bt_set_stream": [P, P],
    "bt_set_params": [P, F, F],
    "bt_set_particles": [P, P, P],
    "bt_set_particles_device": [P, P, P],
    "bt_get_particles": [P, P, P, I],
    "bt_build_tree": [P, I],
    "bt_compute_multipoles": [P],
    "bt_make_groups": [P, I, I],
    "bt_compute_forces": [P, I, P],
    "bt_move_particles": [P, F],
    "bt_kick": [P, F],
    "bt_compute_energies": [P, P, P],
    "bt_step": [P, F, I, I, I],
    "bt_get_acc": [P, P, I],
    "bt_get_potential": [P, P],
    "bt_get_walk_stats": [P, P, P],
    "bt_get_counts": [P, P],
    "bt_get_tree": [P, P, P, P, P, P, P],
    "bt_get_cell_ls": [P, P],
    "bt_get_groups": [P, P],
    "bt_get_group_counts": [P, P],
    "bt_get_ph": [P, P, P],
    "bt_get_ids": [P, P, I],
    "bt_get_boxes": [P, P],
    "bt_direct_fp64": [P, I, P, P],
    "bt_direct_fp64_noalloc": [P, I, P, P],
    "bt_get_timings": [P, P],
    "bt_set_walk_options": [P, I, I, I],
    "bt_set_walk_variant": [P, I],
    "bt_set_walk_lds_stack": [P, I],
    "bt_set_overlap": [P, I],
    "bt_get_walk_stack_max": [P, P],
    "bt_set_group_mode": [P, I],
    "bt_set_quadrupole": [P, I],
    "bt_step_graph": [P, F, I, I, I, I],
    "bt_autotune": [P, I, I, P, P, P],
    "bt_mlp_init": [P, ctypes.c_uint64, F],
    "bt_mlp_set_params": [P, P],
    "bt_mlp_get_params": [P, P],
    "bt_mlp_forward": [P, I, P, P],
    "bt_mlp_grad": [P, I, P, P, P, P],
    "bt_mlp_train": [P, I, P, P, I, I, F, ctypes.c_uint64, P],
    "bt_learn_collect": [P, I, P, F, F, I, P, P, P, P, P, P],
    "bt_set_opening": [P, I, F, F, F],
    "bt_set_opening_check": [P, I, F, ctypes.c_uint64],
    "bt_get_opening_check": [P, P],
    "bt_get_check_samples": [P, P],
    "bt_ic_plummer": [I, ctypes.c_uint64, D, P, P],
    "bt_ic_clustered": [I, I, ctypes.c_uint64, D, D, D, P, P],
    "bt_ic_halo_disk": [I, ctypes.c_uint64, D, D, D, P, P],
    "bt_create_f64": [ctypes.POINTER(P), D, D, I],
    "bt_set_params_f64": [P, D, D],
    "bt_set_particles_f64": [P, P, P],
    "bt_get_particles_f64": [P, P, P, I],
    "bt_get_acc_f64": [P, P, I],
    "bt_get_tree_f64": [P, P, P, P, P],
    "bt_get_boxes_f64": [P, P],
    "bt_move_particles_f64": [P, D],
    "bt_step_f64": [P, D, I, I, I],
    "bt_host_alloc": [ctypes.POINTER(P), ctypes.c_size_t],
    "bt_host_free": [P],
    "bt_get_device_buffer": [P, I, ctypes.POINTER(P), P],
    "bt_copy_device_buffer": [P, I, P, ctypes.c_int64],
    "bt_read_tipsy": [ctypes.c_char_p, P, P, P, I],
    "bt_write_tipsy": [ctypes.c_char_p, I, P, P, D, F],
}
EXPORTED_SYMBOLS = sorted(list(_SIGS) + ["bt_last_error"])

_LIB = None


class BonsaiError(RuntimeError):
    pass


def lib():
    """The loaded libbonsai2_amd.so (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise BonsaiError("libbonsai2_amd.so is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = I
        L.bt_last_error.argtypes = [P]
        L.bt_last_error.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def _f4(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != 4:
        raise ValueError("expected an (n, 4) float32 array")
    return a


def _d4(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != 4:
        raise ValueError("expected an (n, 4) float64 array")
    return a


class Treecode:
    """Drop-in mirror of Treecode<real_t> (code/Treecode.h:164-221).

    Treecode(eps=0.01, theta=0.75, ncrit=64) as the reference constructor;
    fp64=True makes Treecode<double> (code/Particle4.h:106-125): particles,
    moments, walk and integration in FP64, arrays of float64.
    """

    def __init__(self, eps=0.01, theta=0.75, ncrit=64, fp64=False):
        self._L = lib()
        h = P()
        self.fp64 = bool(fp64)
        create = self._L.bt_create_f64 if self.fp64 else self._L.bt_create
        rc = create(ctypes.byref(h), float(eps), float(theta), int(ncrit))
        if rc != 0:
            raise BonsaiError("bt_create failed (%d): no HIP device or runtime error" % rc)
        self._h = h
        self._rt = np.float64 if self.fp64 else np.float32
        self.eps = float(self._rt(eps))
        self.theta = float(self._rt(theta))
        self.nCrit = int(ncrit)
        self.nLeaf = 16
        self.nPtcl = 0
        self.nsys = 1
        self._check_ns = 0
        self.offsets = None
        self.grav_potential = 0.0

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._L.bt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc != 0:
            msg = self._L.bt_last_error(self._h)
            raise BonsaiError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def alloc(self, n, offsets=None):
        """Treecode::alloc(n) (code/Treecode.h:175-198); offsets -> a forest of systems."""
        if offsets is None:
            self._ck(self._L.bt_alloc(self._h, int(n), 1, None), "alloc")
            self.nsys, self.offsets = 1, np.array([0, n], np.int32)
        else:
            off = np.ascontiguousarray(offsets, np.int32)
            self._ck(self._L.bt_alloc(self._h, int(n), len(off) - 1, _p(off)), "alloc")
            self.nsys, self.offsets = len(off) - 1, off
        self.nPtcl = int(n)

    def set_params(self, eps=None, theta=None):
        if eps is not None:
            self.eps = float(self._rt(eps))
        if theta is not None:
            self.theta = float(self._rt(theta))
        fn = self._L.bt_set_params_f64 if self.fp64 else self._L.bt_set_params
        self._ck(fn(self._h, self.eps, self.theta), "set_params")

    def set_walk_options(self, kahan=1, stack_ints=16384, blocks_per_cu=0):
        self._ck(self._L.bt_set_walk_options(self._h, int(kahan), int(stack_ints),
                                             int(blocks_per_cu)), "set_walk_options")

    def set_walk_variant(self, variant):
        """0 = LDS-staged j-data, one i-particle per lane (default); 1 = scalar-load
        j-data; 2 = two i-particles per lane, each half-wave sweeping half a batch."""
        self._ck(self._L.bt_set_walk_variant(self._h, int(variant)), "set_walk_variant")

    def set_walk_lds_stack(self, entries):
        """FP32 walks: stack entries per wave kept in LDS (512..1024)."""
        self._ck(self._L.bt_set_walk_lds_stack(self._h, int(entries)), "set_walk_lds_stack")

    def set_overlap(self, on):
        """Steps run makeGroups' work on a second stream beside the level build
        and multipoles (True, default) or every phase in sequence (False; the
        phase timings are then each phase alone).  Results are bitwise equal."""
        self._ck(self._L.bt_set_overlap(self._h, int(bool(on))), "set_overlap")

    def set_group_mode(self, mode):
        """0 = makeGroups' PH-key groups (default), 1 = one group per leaf, 2 = one
        group per tree cell of <= ncrit particles (both in tree order)."""
        self._ck(self._L.bt_set_group_mode(self._h, int(mode)), "set_group_mode")

    def walk_stack_max(self):
        """Deepest traversal stack (entries) of any group in the last walk."""
        v = np.zeros(1, np.int32)
        self._ck(self._L.bt_get_walk_stack_max(self._h, _p(v)), "walk_stack_max")
        return int(v[0])

    def set_quadrupole(self, on=True):
        """code/Treecode.h:14-16: quadrupole P2C (default) or monopole only."""
        self._ck(self._L.bt_set_quadrupole(self._h, 1 if on else 0), "set_quadrupole")

    # -- captured step and autotuner (no reference counterpart) ---------------
    def step_graph(self, dt, nLeaf=16, levelSplit=5, nCrit=64, nsteps=1):
        """nsteps replays of the step captured in a HIP graph."""
        self._ck(self._L.bt_step_graph(self._h, float(dt), int(nLeaf), int(levelSplit),
                                       int(nCrit), int(nsteps)), "step_graph")

    def autotune(self, levelSplit=5):
        """-> (best dict, table rows) ; the chosen set-up stays active."""
        table = np.zeros((16, 6), np.float64)
        nt = np.zeros(1, np.int32)
        best = np.zeros(4, np.int32)
        self._ck(self._L.bt_autotune(self._h, int(levelSplit), 0, _p(table), _p(nt), _p(best)),
                 "autotune")
        rows = [dict(nleaf=int(r[0]), ncrit=int(r[1]), blocks_per_cu=int(r[2]), step_ms=float(r[3]),
                     walk_ms=float(r[4]), deepest_stack=int(r[5])) for r in table[:int(nt[0])]]
        b = dict(nleaf=int(best[0]), ncrit=int(best[1]), blocks_per_cu=int(best[2]),
                 stack_ints=int(best[3]))
        self.nLeaf, self.nCrit = b["nleaf"], b["ncrit"]
        return b, rows

    # -- learned opening angle (no reference counterpart) --------------------
    MLP_PARAMS = 1377

    def mlp_init(self, seed=1, b3=0.0):
        self._ck(self._L.bt_mlp_init(self._h, int(seed), float(b3)), "mlp_init")

    def set_mlp_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        if p.shape != (self.MLP_PARAMS,):
            raise ValueError("expected 1377 parameters")
        self._ck(self._L.bt_mlp_set_params(self._h, _p(p)), "mlp_set_params")

    def mlp_params(self):
        p = np.empty(self.MLP_PARAMS, np.float32)
        self._ck(self._L.bt_mlp_get_params(self._h, _p(p)), "mlp_get_params")
        return p

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 8)
        y = np.empty(len(x), np.float32)
        self._ck(self._L.bt_mlp_forward(self._h, len(x), _p(x), _p(y)), "mlp_forward")
        return y

    def mlp_grad(self, x, t):
        """(loss, grad[1377]) of the mean squared error on (x, t)."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 8)
        t = np.ascontiguousarray(t, np.float32)
        g = np.empty(self.MLP_PARAMS, np.float64)
        loss = np.zeros(1, np.float64)
        self._ck(self._L.bt_mlp_grad(self._h, len(x), _p(x), _p(t), _p(g), _p(loss)), "mlp_grad")
        return float(loss[0]), g

    def mlp_train(self, x, t, steps=2000, batch=8192, lr=3e-3, seed=1):
        """Adam on the GPU; returns the per-step minibatch loss."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 8)
        t = np.ascontiguousarray(t, np.float32)
        curve = np.empty(int(steps), np.float32)
        self._ck(self._L.bt_mlp_train(self._h, len(x), _p(x), _p(t), int(steps), int(batch),
                                      float(lr), int(seed), _p(curve)), "mlp_train")
        return curve

    def learn_collect(self, gsel, theta_lo=0.25, theta_hi=1.0, max_pairs=1 << 22):
        """Training pairs of the groups gsel -> dict(n, feat, target, err, theta_need, gc)."""
        gsel = np.ascontiguousarray(gsel, np.int32)
        feat = np.empty((max_pairs, 8), np.float32)
        tg = np.empty(max_pairs, np.float32)
        er = np.empty(max_pairs, np.float32)
        tn = np.empty(max_pairs, np.float32)
        gc = np.empty((max_pairs, 2), np.int32)
        n = np.zeros(1, np.int32)
        self._ck(self._L.bt_learn_collect(self._h, len(gsel), _p(gsel), float(theta_lo),
                                          float(theta_hi), int(max_pairs), _p(feat), _p(tg),
                                          _p(er), _p(tn), _p(gc), _p(n)), "learn_collect")
        m = min(int(n[0]), max_pairs)
        return dict(n=int(n[0]), feat=feat[:m], target=tg[:m], err=er[:m], theta_need=tn[:m],
                    gc=gc[:m])

    def set_opening(self, mode, budget=1e-4, theta_lo=0.25, theta_hi=1.0):
        """mode 0: fixed theta; mode 1: the network's per-pair angle."""
        self._ck(self._L.bt_set_opening(self._h, int(mode), float(budget), float(theta_lo),
                                        float(theta_hi)), "set_opening")

    def set_opening_check(self, nsamples=256, p99_max=1e-3, seed=1):
        self._ck(self._L.bt_set_opening_check(self._h, int(nsamples), float(p99_max), int(seed)),
                 "set_opening_check")
        self._check_ns = int(nsamples)

    def use_learned_theta(self, path=None, check_samples=256, p99_max=1e-3, budget=None):
        """Loads a trained network (bonsai2_amd/theta_mlp.json by default) and
        switches the walk to the learned per-pair angle, re-checked against
        check_samples FP64 direct sums after every walk (0 = no check)."""
        from . import learn
        params, b, lo, hi = learn.load(path or learn.DEFAULT_PATH)
        self.set_mlp_params(params)
        self.set_opening(1, b if budget is None else budget, lo, hi)
        self.set_opening_check(check_samples, p99_max)
        return b if budget is None else budget

    def opening_check(self):
        """{p50, p99, max, fell_back, fallbacks, checks} of the last re-check."""
        o = np.zeros(6, np.float64)
        self._ck(self._L.bt_get_opening_check(self._h, _p(o)), "opening_check")
        return dict(p50=o[0], p99=o[1], max=o[2], fell_back=bool(o[3]), fallbacks=int(o[4]),
                    checks=int(o[5]))

    def check_samples(self):
        """PH indices of the last re-check's samples, in draw order."""
        ns = min(self._check_ns, self.nPtcl)
        idx = np.empty(ns, np.int32)
        self._ck(self._L.bt_get_check_samples(self._h, _p(idx)), "check_samples")
        return idx

    # -- particles ---------------------------------------------------------
    def ptcl_h2d(self, pos, vel):
        """ptcl_h2d() (code/Treecode.h:207-214): pos/vel are (n,4) float32
        (float64 for an fp64 context)."""
        cv = _d4 if self.fp64 else _f4
        pos, vel = cv(pos), cv(vel)
        if len(pos) != self.nPtcl or len(vel) != self.nPtcl:
            raise ValueError("particle count differs from alloc()")
        fn = self._L.bt_set_particles_f64 if self.fp64 else self._L.bt_set_particles
        self._ck(fn(self._h, _p(pos), _p(vel)), "ptcl_h2d")

    def ptcl_d2h(self, order=ORDER_TREE):
        """ptcl_d2h() (code/Treecode.h:200-204) -> (pos, vel).  The default is the
        reference's: the device arrays as they are (tree order after
        buildTree, input order before the first build)."""
        pos = np.empty((self.nPtcl, 4), self._rt)
        vel = np.empty((self.nPtcl, 4), self._rt)
        fn = self._L.bt_get_particles_f64 if self.fp64 else self._L.bt_get_particles
        self._ck(fn(self._h, _p(pos), _p(vel), int(order)), "ptcl_d2h")
        return pos, vel

    # -- the gravity step (reference method names) ----------------------------
    def buildTree(self, nLeaf=16):
        self._ck(self._L.bt_build_tree(self._h, int(nLeaf)), "buildTree")
        self.nLeaf = int(nLeaf)

    def computeMultipoles(self):
        self._ck(self._L.bt_compute_multipoles(self._h), "computeMultipoles")

    def makeGroups(self, levelSplit=1, nCrit=64):
        self._ck(self._L.bt_make_groups(self._h, int(levelSplit), int(nCrit)), "makeGroups")
        self.nCrit = int(nCrit)

    def computeForces(self, INTCOUNT=True):
        """Returns (direct <n>, direct max, approx <n>, approx max) (code/computeForces.cu:701-715)."""
        st = np.zeros(4, np.float64)
        self._ck(self._L.bt_compute_forces(self._h, 1 if INTCOUNT else 0, _p(st)),
                 "computeForces")
        pot = np.zeros(1, np.float64)
        self._ck(self._L.bt_get_potential(self._h, _p(pot)), "grav_potential")
        self.grav_potential = float(pot[0])
        return tuple(float(x) for x in st)

    def moveParticles(self, dt):
        """First half of a kick-drift-kick step (v += a dt/2, x += v dt); the
        next force evaluation applies the second half kick."""
        fn = self._L.bt_move_particles_f64 if self.fp64 else self._L.bt_move_particles
        self._ck(fn(self._h, float(dt)), "moveParticles")

    def kick(self, dt_half):
        self._ck(self._L.bt_kick(self._h, float(dt_half)), "kick")

    def computeEnergies(self):
        """FP64 (K, W) per system (arrays of length nsys); evaluates the forces
        first when the particles moved since the last evaluation."""
        ek = np.zeros(self.nsys, np.float64)
        ep = np.zeros(self.nsys, np.float64)
        self._ck(self._L.bt_compute_energies(self._h, _p(ek), _p(ep)), "computeEnergies")
        # the forces (and 1/2 sum m phi) may have been re-evaluated: as the C++ mirror
        pot = np.zeros(1, np.float64)
        self._ck(self._L.bt_get_potential(self._h, _p(pot)), "grav_potential")
        self.grav_potential = float(pot[0])
        return ek, ep

    def step(self, dt, nLeaf=16, levelSplit=5, nCrit=64):
        fn = self._L.bt_step_f64 if self.fp64 else self._L.bt_step
        self._ck(fn(self._h, float(dt), int(nLeaf), int(levelSplit), int(nCrit)), "step")

    # -- results -------------------------------------------------------------
    def get_nPtcl(self):
        return self.nPtcl

    def get_nCrit(self):
        return self.nCrit

    def get_nLeaf(self):
        return self.nLeaf

    def walk_stats(self):
        """(stats[4], totals) of the last walk: totals = (n_P2P, n_P2C) summed over particles."""
        st = np.zeros(4, np.float64)
        tot = np.zeros(2, np.float64)
        self._ck(self._L.bt_get_walk_stats(self._h, _p(st), _p(tot)), "get_walk_stats")
        return st, tot

    def counts(self):
        c = np.zeros(8, np.int32)
        self._ck(self._L.bt_get_counts(self._h, _p(c)), "counts")
        keys = ["n", "ncells", "nnodes", "nleaves", "nlevels", "ngroups", "nsys", "flags"]
        return dict(zip(keys, (int(x) for x in c)))

    def acc(self, order=ORDER_PH):
        a = np.empty((self.nPtcl, 4), self._rt)
        fn = self._L.bt_get_acc_f64 if self.fp64 else self._L.bt_get_acc
        self._ck(fn(self._h, _p(a), int(order)), "get_acc")
        return a

    def tree(self, multipoles=True):
        """Cells (CellData uint4), level table int2[32] and, optionally, the moments."""
        nc = self.counts()["ncells"]
        cells = np.empty((nc, 4), np.uint32)
        lr = np.empty((32, 2), np.int32)
        if multipoles:
            size = np.empty((nc, 4), self._rt)
            mono = np.empty((nc, 4), self._rt)
            q0 = np.empty((nc, 4), self._rt)
            q1 = np.empty((nc, 2), self._rt)
            if self.fp64:
                self._ck(self._L.bt_get_tree(self._h, _p(cells), _p(lr), None, None, None, None),
                         "get_tree")
                self._ck(self._L.bt_get_tree_f64(self._h, _p(size), _p(mono), _p(q0), _p(q1)),
                         "get_tree_f64")
            else:
                self._ck(self._L.bt_get_tree(self._h, _p(cells), _p(lr), _p(size), _p(mono),
                                             _p(q0), _p(q1)), "get_tree")
            return dict(cells=cells, level_range=lr, size=size, mono=mono, q0=q0, q1=q1)
        self._ck(self._L.bt_get_tree(self._h, _p(cells), _p(lr), None, None, None, None),
                 "get_tree")
        return dict(cells=cells, level_range=lr)

    def cell_ls(self):
        nc = self.counts()["ncells"]
        ls = np.empty((nc, 2), np.float32)
        self._ck(self._L.bt_get_cell_ls(self._h, _p(ls)), "get_cell_ls")
        return ls

    def groups(self):
        ng = self.counts()["ngroups"]
        g = np.empty((ng, 2), np.int32)
        self._ck(self._L.bt_get_groups(self._h, _p(g)), "get_groups")
        return g

    def group_counts(self):
        ng = self.counts()["ngroups"]
        g = np.empty((ng, 2), np.int32)
        self._ck(self._L.bt_get_group_counts(self._h, _p(g)), "get_group_counts")
        return g

    def ph(self):
        keys = np.empty(self.nPtcl, np.uint64)
        perm = np.empty(self.nPtcl, np.int32)
        self._ck(self._L.bt_get_ph(self._h, _p(keys), _p(perm)), "get_ph")
        return keys, perm

    def ids(self, order):
        ids = np.empty(self.nPtcl, np.int32)
        self._ck(self._L.bt_get_ids(self._h, _p(ids), int(order)), "get_ids")
        return ids

    def boxes(self):
        b = np.empty((self.nsys, 4), self._rt)
        fn = self._L.bt_get_boxes_f64 if self.fp64 else self._L.bt_get_boxes
        self._ck(fn(self._h, _p(b)), "get_boxes")
        return b

    def direct_fp64(self, ph_idx):
        idx = np.ascontiguousarray(ph_idx, np.int32)
        out = np.empty((len(idx), 4), np.float64)
        self._ck(self._L.bt_direct_fp64(self._h, len(idx), _p(idx), _p(out)), "direct_fp64")
        return out

    def direct_fp64_noalloc(self, ns, d_ph_idx, d_out4):
        """bt_direct_fp64_noalloc: device pointers in and out (ints) on the
        context stream, no allocation and no synchronisation; returns the
        status code (0 = ok) without raising."""
        return int(self._L.bt_direct_fp64_noalloc(self._h, int(ns), P(d_ph_idx), P(d_out4)))

    def timings(self):
        t = np.zeros(10, np.float32)
        self._ck(self._L.bt_get_timings(self._h, _p(t)), "get_timings")
        names = ["bbox_keys", "sort", "levels", "multipoles", "ph_keys", "ph_sort", "groups",
                 "walk", "integrate", "energies"]
        return dict(zip(names, (float(x) for x in t)))


# ---- initial conditions and tipsy I/O (host only) ---------------------------

def _ic(fn, n, *args):
    pos = np.empty((n, 4), np.float32)
    vel = np.empty((n, 4), np.float32)
    rc = fn(int(n), *args, _p(pos), _p(vel))
    if rc != 0:
        raise BonsaiError("initial-condition generator failed (%d)" % rc)
    return pos, vel


def plummer(n, seed=19810614, rmax=22.8):
    """Plummer sphere in Henon units, truncated at rmax (structural units)."""
    return _ic(lib().bt_ic_plummer, n, ctypes.c_uint64(seed), float(rmax))


def clustered(n, nclust=64, seed=19810614, box=10.0, mmin=1e-3, a0=0.5):
    return _ic(lib().bt_ic_clustered, n, int(nclust), ctypes.c_uint64(seed), float(box),
               float(mmin), float(a0))


def halo_disk(n, seed=19810614, mh=0.8, rd=0.3, z0=0.03):
    return _ic(lib().bt_ic_halo_disk, n, ctypes.c_uint64(seed), float(mh), float(rd), float(z0))


def ensemble(nsys=256, seed=19810614, log2_min=10.0, log2_max=16.0, rmax=22.8):
    """Independent Plummer systems (Henon units each, mass 1 per system) with
    N_k = round(2^U[log2_min, log2_max]) from a seeded draw -> (pos, vel,
    offsets) for Treecode.alloc(n, offsets)."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(2.0 ** rng.uniform(log2_min, log2_max, nsys)).astype(np.int64)
    parts = [plummer(int(nk), seed=seed + 1 + k, rmax=rmax) for k, nk in enumerate(sizes)]
    pos = np.concatenate([p for p, _ in parts])
    vel = np.concatenate([v for _, v in parts])
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    return pos, vel, offsets


def read_tipsy(path):
    n = np.zeros(1, np.int32)
    rc = lib().bt_read_tipsy(path.encode(), _p(n), None, None, 0)
    if rc != 0:
        raise BonsaiError("read_tipsy failed (%d)" % rc)
    pos = np.empty((int(n[0]), 4), np.float32)
    vel = np.empty((int(n[0]), 4), np.float32)
    rc = lib().bt_read_tipsy(path.encode(), _p(n), _p(pos), _p(vel), int(n[0]))
    if rc != 0:
        raise BonsaiError("read_tipsy failed (%d)" % rc)
    return pos, vel


def write_tipsy(path, pos, vel, time=0.0, eps=0.0):
    pos, vel = _f4(pos), _f4(vel)
    rc = lib().bt_write_tipsy(path.encode(), len(pos), _p(pos), _p(vel), float(time), float(eps))
    if rc != 0:
        raise BonsaiError("write_tipsy failed (%d)" % rc)
