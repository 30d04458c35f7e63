"""numpy/ctypes front end of the CPU oracle (liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The product never
imports this module.  Each wrapper names the reference file:line that the
C function restates (see oracle.cpp for the full citations).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

P = ctypes.c_void_p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(P)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.or_version.restype = ctypes.c_int
        L.or_bounding_box.argtypes = [ctypes.c_int, P, P, P, P]
        L.or_geo_keys.argtypes = [ctypes.c_int, P, P, P]
        L.or_build_tree.argtypes = [ctypes.c_int, P, ctypes.c_int, P, P, ctypes.c_int,
                                    ctypes.c_int, P, P, P, P, P]
        L.or_build_tree.restype = ctypes.c_int
        L.or_multipoles.argtypes = [ctypes.c_int, P, P, ctypes.c_float, ctypes.c_int,
                                    P, P, P, P, P]
        L.or_hilbert_key.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_hilbert_key.restype = ctypes.c_uint64
        L.or_ph_keys.argtypes = [ctypes.c_int, P, P, P]
        L.or_sort_keys.argtypes = [ctypes.c_int, P, P]
        L.or_make_groups.argtypes = [ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int, P]
        L.or_make_groups.restype = ctypes.c_int
        L.or_walk.argtypes = [ctypes.c_int, P, P, P, P, P, P, P, P, P, P, ctypes.c_float,
                              ctypes.c_int, P, P, P, P, P, P, ctypes.c_int]
        L.or_bounding_box_d.argtypes = [ctypes.c_int, P, P, P, P]
        L.or_geo_keys_d.argtypes = [ctypes.c_int, P, P, P]
        L.or_build_tree_d.argtypes = [ctypes.c_int, P, ctypes.c_int, P, P, ctypes.c_int,
                                      ctypes.c_int, P, P, P, P, P]
        L.or_build_tree_d.restype = ctypes.c_int
        L.or_multipoles_d.argtypes = [ctypes.c_int, P, P, ctypes.c_double, P, P, P, P, P]
        L.or_ph_keys_d.argtypes = [ctypes.c_int, P, P, P]
        L.or_walk_d.argtypes = [ctypes.c_int, P, P, P, P, P, P, P, P, P, ctypes.c_double, P, P,
                                P, P, ctypes.c_int]
        L.or_direct_fp64_d.argtypes = [ctypes.c_int, P, ctypes.c_int, P, ctypes.c_double, P]
        L.or_walk_lists.argtypes = [ctypes.c_int, P, P, P, P, P, P, ctypes.c_float,
                                    ctypes.c_int, P, P, P, P]
        L.or_direct_fp64.argtypes = [ctypes.c_int, P, ctypes.c_int, P, ctypes.c_double, P]
        L.or_direct_ld.argtypes = [ctypes.c_int, P, ctypes.c_int, P, ctypes.c_double, P]
        L.or_energies.argtypes = [ctypes.c_int, P, P, P, ctypes.c_double, P, P]
        L.or_plummer_ref.argtypes = [ctypes.c_long, ctypes.c_uint, ctypes.c_double, P, P, P]
        L.or_f16_round.argtypes = [ctypes.c_int, P, P]
        L.or_pair_features.argtypes = [ctypes.c_int, P, P]
        L.or_mlp_forward.argtypes = [ctypes.c_int, P, P, ctypes.c_int, P]
        L.or_mlp_grad.argtypes = [ctypes.c_int, P, P, P, ctypes.c_int, P, P]
        L.or_walk_learned.argtypes = [ctypes.c_int, P, P, P, P, P, P, P, P, P, P, P,
                                      ctypes.c_float, P, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, P, P, P, P, P, P]
        L.or_learn_pairs.argtypes = [ctypes.c_int, P, P, P, P, P, P, P, P, P, P, P, P, P,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_float, P,
                                     ctypes.c_int, P, P, P, P, P]
        L.or_learn_pairs.restype = ctypes.c_int
        _LIB = L
    return _LIB


def ref_host():
    """oracle/_ref/libref_host.so (the reference's own plummer.h/read_tipsy.h), or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_host.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.ref_plummer.argtypes = [ctypes.c_long, ctypes.c_uint, P, P, P]
        R.ref_plummer.restype = ctypes.c_int
        R.ref_read_tipsy.argtypes = [ctypes.c_char_p, ctypes.c_int, P, P, P]
        R.ref_read_tipsy.restype = ctypes.c_int
        _REF = R
    return _REF


def _f4(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    assert a.ndim == 2 and a.shape[1] == 4
    return a


def _d4(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    assert a.ndim == 2 and a.shape[1] == 4
    return a


def _is64(pos):
    return np.asarray(pos).dtype == np.float64


def bounding_box(pos):
    """code/buildTree.cu:80-166 -> (box[4] = cx,cy,cz,hsize, bmin[3], bmax[3]).
    float64 positions: the Treecode<double> box (extremes kept in FP64)."""
    if _is64(pos):
        pos = _d4(pos)
        box, bmin, bmax = np.zeros(4), np.zeros(3), np.zeros(3)
        lib().or_bounding_box_d(len(pos), _ptr(pos), _ptr(box), _ptr(bmin), _ptr(bmax))
        return box, bmin, bmax
    pos = _f4(pos)
    box = np.zeros(4, np.float32)
    bmin = np.zeros(3, np.float32)
    bmax = np.zeros(3, np.float32)
    lib().or_bounding_box(len(pos), _ptr(pos), _ptr(box), _ptr(bmin), _ptr(bmax))
    return box, bmin, bmax


def geo_keys(pos, box):
    """Octant digits down the ChildBox recursion (code/buildTree.cu:233-247)."""
    keys = np.zeros(len(pos), np.uint64)
    if _is64(pos):
        lib().or_geo_keys_d(len(pos), _ptr(_d4(pos)), _ptr(np.ascontiguousarray(box, np.float64)),
                            _ptr(keys))
        return keys
    pos = _f4(pos)
    box = np.ascontiguousarray(box, np.float32)
    lib().or_geo_keys(len(pos), _ptr(pos), _ptr(box), _ptr(keys))
    return keys


class Tree:
    """Result of or_build_tree (code/buildTree.cu:170-804, canonical order)."""

    def __init__(self, order, cells, level_range, root_cells, stats):
        self.order = order
        self.cells = cells
        self.level_range = level_range
        self.root_cells = root_cells
        self.ncells, self.nnodes, self.nleaves, self.nlevels, self.capped, self.nb_leaf = (
            int(x) for x in stats)


def build_tree(pos, nleaf, box=None, offsets=None, boxes=None):
    fp64 = _is64(pos)
    pos = _d4(pos) if fp64 else _f4(pos)
    rt = np.float64 if fp64 else np.float32
    n = len(pos)
    if offsets is None:
        offsets = np.array([0, n], np.int32)
        if box is None:
            box = bounding_box(pos)[0]
        boxes = np.ascontiguousarray(box, rt).reshape(1, 4)
    offsets = np.ascontiguousarray(offsets, np.int32)
    boxes = np.ascontiguousarray(boxes, rt)
    nroot = len(offsets) - 1
    ncell_max = 3 * n + 64 * nroot + 1024   # the largest tree (levels add <= n cells each)
    order = np.zeros(n, np.int32)
    cells = np.zeros((ncell_max, 4), np.uint32)
    level_range = np.zeros(64, np.int32)
    root_cells = np.zeros(2 * nroot, np.int32)
    stats = np.zeros(6, np.int32)
    fn = lib().or_build_tree_d if fp64 else lib().or_build_tree
    rc = fn(n, _ptr(pos), nroot, _ptr(offsets), _ptr(boxes), int(nleaf), ncell_max, _ptr(order),
            _ptr(cells), _ptr(level_range), _ptr(root_cells), _ptr(stats))
    if rc != 0:
        raise RuntimeError("or_build_tree failed: %d" % rc)
    t = Tree(order, cells, level_range.reshape(32, 2), root_cells.reshape(nroot, 2), stats)
    t.cells = cells[: t.ncells].copy()
    return t


def multipoles(cells, pos_tree, theta, mode=0):
    """code/computeMultipoles.cu:81-190 -> size, mono, quad0, quad1, ls."""
    cells = np.ascontiguousarray(cells, np.uint32)
    pos_tree = _f4(pos_tree)
    nc = len(cells)
    inv_theta = np.float32(1.0 / float(np.float32(theta)))
    size = np.zeros((nc, 4), np.float32)
    mono = np.zeros((nc, 4), np.float32)
    q0 = np.zeros((nc, 4), np.float32)
    q1 = np.zeros((nc, 2), np.float32)
    ls = np.zeros((nc, 2), np.float32)
    lib().or_multipoles(nc, _ptr(cells), _ptr(pos_tree), ctypes.c_float(inv_theta), int(mode),
                        _ptr(size), _ptr(mono), _ptr(q0), _ptr(q1), _ptr(ls))
    return size, mono, q0, q1, ls


def multipoles_d(cells, pos_tree, theta):
    """Treecode<double>: code/computeMultipoles.cu:81-190 in FP64 throughout ->
    size, mono, quad0, quad1, ls (float64)."""
    cells = np.ascontiguousarray(cells, np.uint32)
    pos_tree = _d4(pos_tree)
    nc = len(cells)
    size, mono, q0 = np.zeros((nc, 4)), np.zeros((nc, 4)), np.zeros((nc, 4))
    q1, ls = np.zeros((nc, 2)), np.zeros((nc, 2))
    lib().or_multipoles_d(nc, _ptr(cells), _ptr(pos_tree), ctypes.c_double(1.0 / float(theta)),
                          _ptr(size), _ptr(mono), _ptr(q0), _ptr(q1), _ptr(ls))
    return size, mono, q0, q1, ls


def hilbert_key(x, y, z, nbits=21):
    return int(lib().or_hilbert_key(int(x), int(y), int(z), int(nbits)))


def ph_keys(pos, box):
    """code/makeGroups.cu:19-135."""
    if _is64(pos):
        keys = np.zeros(len(pos), np.uint64)
        lib().or_ph_keys_d(len(pos), _ptr(_d4(pos)), _ptr(np.ascontiguousarray(box, np.float64)),
                           _ptr(keys))
        return keys
    pos = _f4(pos)
    box = np.ascontiguousarray(box, np.float32)
    keys = np.zeros(len(pos), np.uint64)
    lib().or_ph_keys(len(pos), _ptr(pos), _ptr(box), _ptr(keys))
    return keys


def sort_keys(keys):
    """Stable sort permutation (code/makeGroups.cu:270-277 contract)."""
    keys = np.ascontiguousarray(keys, np.uint64)
    perm = np.zeros(len(keys), np.int32)
    lib().or_sort_keys(len(keys), _ptr(keys), _ptr(perm))
    return perm


def make_groups(sorted_keys, level_split, ncrit, seg_id=None):
    """code/makeGroups.cu:137-223 -> int32 (ngroups, 2) of (pbeg, np)."""
    sorted_keys = np.ascontiguousarray(sorted_keys, np.uint64)
    n = len(sorted_keys)
    out = np.zeros((max(n, 1), 2), np.int32)
    if seg_id is not None:
        seg_id = np.ascontiguousarray(seg_id, np.int32)
    ng = lib().or_make_groups(n, _ptr(sorted_keys), _ptr(seg_id), int(level_split), int(ncrit),
                              _ptr(out))
    return out[:ng].copy()


def leaf_groups(cells):
    """One group per leaf cell, ordered by first particle (tree order): the
    leaf list of code/buildTree.cu:954-958 (collect_leaves, :807-850; unused
    by the reference, whose list order comes from an atomic counter) as
    groups.  -> int32 (nleaves, 2) of (pbeg, np)."""
    cells = np.asarray(cells, np.uint32).reshape(-1, 4)
    leaf = cells[:, 1] == 0xFFFFFFFF
    beg = cells[leaf, 2].astype(np.int64)
    end = cells[leaf, 3].astype(np.int64)
    order = np.argsort(beg, kind="stable")
    return np.stack([beg[order], end[order] - beg[order]], axis=1).astype(np.int32)


def tree_groups(cells, ncrit):
    """One group per tree cell of at most ncrit particles whose parent holds
    more (level-1 cells have no parent cell), and per leaf of more than ncrit,
    ordered by first particle: the groups the octree itself defines.
    -> int32 (ngroups, 2) of (pbeg, np)."""
    cells = np.asarray(cells, np.uint32).reshape(-1, 4)
    leaf = cells[:, 1] == 0xFFFFFFFF
    np_ = cells[:, 3].astype(np.int64) - cells[:, 2].astype(np.int64)
    level = cells[:, 0] >> 27
    parent = (cells[:, 0] & 0x07FFFFFF).astype(np.int64)
    parent_np = np.where(level > 1, np_[parent], ncrit + 1)
    group = (leaf | (np_ <= ncrit)) & (parent_np > ncrit)
    beg = cells[group, 2].astype(np.int64)
    order = np.argsort(beg, kind="stable")
    return np.stack([beg[order], np_[group][order]], axis=1).astype(np.int32)


def walk(groups, level_range, pos_ph, pos_tree, cells, size, mono, q0, q1, eps,
         acc_type=0, group_cells=None, cell_ls=None, group_inv_theta=None, counts=False,
         monopole=False):
    """code/computeForces.cu:30-727 -> (acc[N,4] PH order, stats[4], pot[, counts]).
    monopole: P2C with the cell's monopole only (the reference without
    QUADRUPOLE, code/computeForces.cu:198-225)."""
    groups = np.ascontiguousarray(groups, np.int32)
    ng = len(groups)
    level_range = np.ascontiguousarray(level_range, np.int32).reshape(-1)
    pos_ph = _f4(pos_ph)
    pos_tree = _f4(pos_tree)
    acc = np.zeros((len(pos_ph), 4), np.float32)
    stats = np.zeros(4, np.int64)
    pot = np.zeros(1, np.float64)
    cnt = np.zeros((ng, 2), np.int32) if counts else None
    if group_cells is not None:
        group_cells = np.ascontiguousarray(group_cells, np.int32)
    if cell_ls is not None:
        cell_ls = np.ascontiguousarray(cell_ls, np.float32)
    if group_inv_theta is not None:
        group_inv_theta = np.ascontiguousarray(group_inv_theta, np.float32)
    eps2 = np.float32(np.float32(eps) * np.float32(eps))
    lib().or_walk(ng, _ptr(groups), _ptr(group_cells), _ptr(level_range), _ptr(pos_ph),
                  _ptr(pos_tree), _ptr(np.ascontiguousarray(cells, np.uint32)),
                  _ptr(np.ascontiguousarray(size, np.float32)),
                  _ptr(np.ascontiguousarray(mono, np.float32)),
                  _ptr(np.ascontiguousarray(q0, np.float32)),
                  _ptr(np.ascontiguousarray(q1, np.float32)), ctypes.c_float(eps2),
                  int(acc_type), _ptr(cell_ls), _ptr(group_inv_theta), _ptr(acc), _ptr(stats),
                  _ptr(pot), _ptr(cnt), int(bool(monopole)))
    if counts:
        return acc, stats, float(pot[0]), cnt
    return acc, stats, float(pot[0])


def walk_lists(g, groups, level_range, pos_ph, cells, size, ls=None, theta=None, mutate=0):
    """Interaction lists of group g (code/computeForces.cu:229-463) ->
    (P2C cells, leaves opened for P2P).  mutate plants a known mistake in the
    opening rule (1: no < 3 particles rule, 2: group point instead of box,
    3: no s term); used only to show the tests catch it."""
    cells = np.ascontiguousarray(cells, np.uint32)
    nc = len(cells)
    alist = np.zeros(nc, np.int32)
    leaves = np.zeros(nc, np.int32)
    na = np.zeros(1, np.int32)
    nl = np.zeros(1, np.int32)
    it = np.float32(1.0 / float(np.float32(theta))) if theta is not None else np.float32(0)
    lib().or_walk_lists(int(g), _ptr(np.ascontiguousarray(groups, np.int32)),
                        _ptr(np.ascontiguousarray(level_range, np.int32).reshape(-1)),
                        _ptr(_f4(pos_ph)), _ptr(cells), _ptr(np.ascontiguousarray(size, np.float32)),
                        _ptr(None if ls is None else np.ascontiguousarray(ls, np.float32)),
                        ctypes.c_float(it), int(mutate), _ptr(alist), _ptr(na), _ptr(leaves),
                        _ptr(nl))
    return alist[:na[0]].copy(), leaves[:nl[0]].copy()


def walk_d(groups, level_range, pos_ph, pos_tree, cells, size, mono, q0, q1, eps,
           monopole=False):
    """Treecode<double>: the walk of code/computeForces.cu:30-727 in FP64 ->
    (acc[N,4] float64 PH order, stats[4], pot, per-group counts)."""
    groups = np.ascontiguousarray(groups, np.int32)
    ng = len(groups)
    acc = np.zeros((len(pos_ph), 4))
    stats = np.zeros(4, np.int64)
    pot = np.zeros(1)
    cnt = np.zeros((ng, 2), np.int32)
    lib().or_walk_d(ng, _ptr(groups), _ptr(np.ascontiguousarray(level_range, np.int32).reshape(-1)),
                    _ptr(_d4(pos_ph)), _ptr(_d4(pos_tree)),
                    _ptr(np.ascontiguousarray(cells, np.uint32)), _ptr(_d4(size)), _ptr(_d4(mono)),
                    _ptr(_d4(q0)), _ptr(np.ascontiguousarray(q1, np.float64)),
                    ctypes.c_double(float(eps) ** 2), _ptr(acc), _ptr(stats), _ptr(pot), _ptr(cnt),
                    int(bool(monopole)))
    return acc, stats, float(pot[0]), cnt


def reference_pipeline_d(pos, nleaf, ncrit, theta, eps, level_split=5):
    """Treecode<double> through the reference call order on the oracle (FP64
    positions, box, keys, moments and walk)."""
    pos = _d4(pos)
    box, _, _ = bounding_box(pos)
    tree = build_tree(pos, nleaf, box)
    pos_tree = pos[tree.order]
    size, mono, q0, q1, ls = multipoles_d(tree.cells, pos_tree, theta)
    keys = ph_keys(pos_tree, box)
    perm = sort_keys(keys)
    skeys = keys[perm]
    pos_ph = pos_tree[perm]
    groups = make_groups(skeys, level_split, ncrit)
    acc, stats, pot, cnt = walk_d(groups, tree.level_range, pos_ph, pos_tree, tree.cells, size,
                                  mono, q0, q1, eps)
    return dict(box=box, tree=tree, pos_tree=pos_tree, size=size, mono=mono, q0=q0, q1=q1,
                ls=ls, keys=keys, perm=perm, sorted_keys=skeys, pos_ph=pos_ph, groups=groups,
                acc=acc, stats=stats, pot=pot, counts=cnt)


def direct_fp64(pos, idx, eps, longdouble=False):
    """FP64 direct sum over all sources for targets idx -> float64 (ns, 4).
    float64 positions: sources and targets in FP64 (Treecode<double>)."""
    if _is64(pos):
        pos = _d4(pos)
        idx = np.ascontiguousarray(idx, np.int32)
        out = np.zeros((len(idx), 4), np.float64)
        lib().or_direct_fp64_d(len(pos), _ptr(pos), len(idx), _ptr(idx),
                               ctypes.c_double(float(eps) ** 2), _ptr(out))
        return out
    pos = _f4(pos)
    idx = np.ascontiguousarray(idx, np.int32)
    out = np.zeros((len(idx), 4), np.float64)
    fn = lib().or_direct_ld if longdouble else lib().or_direct_fp64
    e = float(np.float32(eps))  # the softening the float32 step uses
    fn(len(pos), _ptr(pos), len(idx), _ptr(idx), ctypes.c_double(e * e), _ptr(out))
    return out


def energies(pos, vel, acc, eps):
    ek = np.zeros(1, np.float64)
    ep = np.zeros(1, np.float64)
    lib().or_energies(len(pos), _ptr(_f4(pos)), _ptr(_f4(vel)), _ptr(_f4(acc)),
                      ctypes.c_double(float(np.float32(eps))), _ptr(ek), _ptr(ep))
    return float(ek[0]), float(ep[0])


def plummer_ref(n, seed=19810614, rmax=100.0):
    """code/plummer.h:9-137 restated (glibc rand) -> mass[n], pos[n,3], vel[n,3] (float64)."""
    mass = np.zeros(n, np.float64)
    pos = np.zeros((n, 3), np.float64)
    vel = np.zeros((n, 3), np.float64)
    lib().or_plummer_ref(int(n), int(seed), float(rmax), _ptr(mass), _ptr(pos), _ptr(vel))
    return mass, pos, vel


def plummer_ref_f4(n, seed=19810614, rmax=100.0):
    """Reference driver's float packing (code/treecode.cu:68-87): pos.w = vel.w = mass."""
    mass, pos, vel = plummer_ref(n, seed, rmax)
    p = np.empty((n, 4), np.float32)
    v = np.empty((n, 4), np.float32)
    p[:, :3] = pos
    p[:, 3] = mass
    v[:, :3] = vel
    v[:, 3] = mass
    return p, v


def reference_pipeline(pos, nleaf, ncrit, theta, eps, level_split=5, acc_type=0, mode=0):
    """The reference call order (code/treecode.cu:137-147) on the oracle:
    buildTree -> computeMultipoles -> makeGroups -> computeForces.
    Returns a dict with every intermediate, all in the reference layouts."""
    pos = _f4(pos)
    box, _, _ = bounding_box(pos)
    tree = build_tree(pos, nleaf, box)
    pos_tree = pos[tree.order]
    size, mono, q0, q1, ls = multipoles(tree.cells, pos_tree, theta, mode)
    keys = ph_keys(pos_tree, box)
    perm = sort_keys(keys)
    skeys = keys[perm]
    pos_ph = pos_tree[perm]
    groups = make_groups(skeys, level_split, ncrit)
    acc, stats, pot = walk(groups, tree.level_range, pos_ph, pos_tree, tree.cells, size, mono,
                           q0, q1, eps, acc_type)
    return dict(box=box, tree=tree, pos_tree=pos_tree, size=size, mono=mono, q0=q0, q1=q1,
                ls=ls, keys=keys, perm=perm, sorted_keys=skeys, pos_ph=pos_ph, groups=groups,
                acc=acc, stats=stats, pot=pot)


# ---------------------------------------------------------------------------
# Learned opening angle (no reference counterpart): features, the 8-32-32-1
# network with FP16 parameters, its gradient, the walk with per-pair angles
# and the training-pair generator.  Parameter vector: 1377 float32 in the
# order W1[32][8], b1[32], W2[32][32], b2[32], w3[32], b3.
# ---------------------------------------------------------------------------
MLP_PARAMS = 1377


def f16_round(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().or_f16_round(x.size, _ptr(x), _ptr(out))
    return out


def pair_features(pin):
    """pin: (n, 20) float32 rows gc.xyz, gs.xyz, np, com.xyz, M, l, Qxx, Qyy,
    Qzz, Qxy, Qxz, Qyz, R, M_sys -> (n, 8) features."""
    pin = np.ascontiguousarray(pin, np.float32)
    out = np.empty((len(pin), 8), np.float32)
    lib().or_pair_features(len(pin), _ptr(pin), _ptr(out))
    return out


def mlp_forward(params, x, round16=True):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty(len(x), np.float32)
    lib().or_mlp_forward(len(x), _ptr(params), _ptr(x), int(round16), _ptr(y))
    return y


def mlp_grad(params, x, t, round16=True):
    """(loss, grad[1377]) of the mean squared error, FP64."""
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    g = np.zeros(MLP_PARAMS, np.float64)
    loss = np.zeros(1, np.float64)
    lib().or_mlp_grad(len(x), _ptr(params), _ptr(x), _ptr(t), int(round16), _ptr(g), _ptr(loss))
    return float(loss[0]), g


def group_scale(groups, R, msys):
    """Per-group (R, M_sys) of a single system."""
    gs = np.empty((len(groups), 2), np.float32)
    gs[:, 0] = R
    gs[:, 1] = msys
    return gs


def walk_learned(groups, level_range, pos_ph, pos_tree, cells, size, mono, q0, q1, ls, eps,
                 params, shift, theta_lo, theta_hi, gscale, group_cells=None, visits=None):
    """The walk with the learned per-pair angle -> (acc, stats, pot, counts).
    visits (int64[2], optional) receives {cells tested, cells whose decision
    depended on the network}."""
    groups = np.ascontiguousarray(groups, np.int32)
    ng = len(groups)
    acc = np.zeros((len(pos_ph), 4), np.float32)
    stats = np.zeros(4, np.int64)
    pot = np.zeros(1, np.float64)
    cnt = np.zeros((ng, 2), np.int32)
    if group_cells is not None:
        group_cells = np.ascontiguousarray(group_cells, np.int32)
    eps2 = np.float32(np.float32(eps) * np.float32(eps))
    lib().or_walk_learned(
        ng, _ptr(groups), _ptr(group_cells),
        _ptr(np.ascontiguousarray(level_range, np.int32).reshape(-1)), _ptr(_f4(pos_ph)),
        _ptr(_f4(pos_tree)), _ptr(np.ascontiguousarray(cells, np.uint32)),
        _ptr(np.ascontiguousarray(size, np.float32)), _ptr(np.ascontiguousarray(mono, np.float32)),
        _ptr(np.ascontiguousarray(q0, np.float32)), _ptr(np.ascontiguousarray(q1, np.float32)),
        _ptr(np.ascontiguousarray(ls, np.float32)), ctypes.c_float(eps2),
        _ptr(np.ascontiguousarray(params, np.float32)), ctypes.c_float(shift),
        ctypes.c_float(theta_lo), ctypes.c_float(theta_hi),
        _ptr(np.ascontiguousarray(gscale, np.float32)), _ptr(acc), _ptr(stats), _ptr(pot),
        _ptr(cnt), _ptr(visits))
    return acc, stats, float(pot[0]), cnt


def learn_pairs(gsel, groups, level_range, pos_ph, pos_tree, cells, size, mono, q0, q1, ls,
                gscale, theta_lo, theta_hi, eps, amag, max_pairs=1 << 20, group_cells=None):
    """Training pairs -> dict(gc[n,2], feat[n,8], theta_need, target, err)."""
    gsel = np.ascontiguousarray(gsel, np.int32)
    gc = np.zeros((max_pairs, 2), np.int32)
    feat = np.zeros((max_pairs, 8), np.float32)
    tn = np.zeros(max_pairs, np.float32)
    tg = np.zeros(max_pairs, np.float32)
    er = np.zeros(max_pairs, np.float32)
    if group_cells is not None:
        group_cells = np.ascontiguousarray(group_cells, np.int32)
    eps2 = np.float32(np.float32(eps) * np.float32(eps))
    n = lib().or_learn_pairs(
        len(gsel), _ptr(gsel), _ptr(np.ascontiguousarray(groups, np.int32)), _ptr(group_cells),
        _ptr(np.ascontiguousarray(level_range, np.int32).reshape(-1)), _ptr(_f4(pos_ph)),
        _ptr(_f4(pos_tree)), _ptr(np.ascontiguousarray(cells, np.uint32)),
        _ptr(np.ascontiguousarray(size, np.float32)), _ptr(np.ascontiguousarray(mono, np.float32)),
        _ptr(np.ascontiguousarray(q0, np.float32)), _ptr(np.ascontiguousarray(q1, np.float32)),
        _ptr(np.ascontiguousarray(ls, np.float32)), _ptr(np.ascontiguousarray(gscale, np.float32)),
        ctypes.c_float(theta_lo), ctypes.c_float(theta_hi), ctypes.c_float(eps2),
        _ptr(np.ascontiguousarray(amag, np.float64)), int(max_pairs), _ptr(gc), _ptr(feat),
        _ptr(tn), _ptr(tg), _ptr(er))
    m = min(n, max_pairs)
    return dict(n=n, gc=gc[:m], feat=feat[:m], theta_need=tn[:m], target=tg[:m], err=er[:m])
