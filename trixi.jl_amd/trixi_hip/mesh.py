"""Uniform TreeMesh, periodic or bounded (host-side setup, no device work).

Restates the parts of Trixi.jl that fix element order and geometry:
  TreeMesh(coordinates_min, coordinates_max; initial_refinement_level, periodicity)
      src/meshes/tree_mesh.jl:128-185
  refine_uniformly! / init_children! (depth-first, children x fastest)
      src/meshes/abstract_tree.jl:206-250, child_sign table :118-128,
      child_coordinates :668-674; leaf order = ascending cell id :154-166
  init_neighbors! with the periodic level-0 cell its own neighbour, and no
      neighbour across a non-periodic domain face
      abstract_tree.jl:252-313, serial_tree.jl:101-140
  init_elements! (inverse_jacobian, node_coordinates)
      src/solvers/dgsem_tree/containers_3d.jl:87-128 (2D: containers_2d.jl)
  init_interfaces! (per element, positive directions, left = element)
      containers_3d.jl:229-279, containers_2d.jl:228-290
  init_boundaries! (per direction, then element order; neighbor_sides 1 when
      the element is left of the boundary) containers_2d.jl:475-543
  refinement_patches of type "box": refine_box! / refine! / rebalance!
      abstract_tree.jl:317-490 (children inserted behind their parent, 2:1
      balance); init_mortars! containers_3d.jl:652-788, containers_2d.jl:700-800
  AMR adaptation of the tree (refine! / rebalance! / coarsen!,
      abstract_tree.jl:367-600; the AMRCallback's selection, amr.jl:244-392)

Refinement, coarsening and the connectivity of refined trees run in native
code (trixi_hip_tree_*, csrc/tree_host.cpp).
"""
import ctypes

import numpy as np

from . import _lib


class TreeMesh:
    def __init__(self, coordinates_min, coordinates_max, initial_refinement_level,
                 periodicity=False, refinement_patches=()):
        cmin = tuple(float(c) for c in coordinates_min)
        cmax = tuple(float(c) for c in coordinates_max)
        if len(cmin) not in (2, 3) or len(cmin) != len(cmax):
            raise ValueError("TreeMesh supports 2D and 3D here")
        if not (isinstance(initial_refinement_level, int) and initial_refinement_level >= 0):
            raise ValueError("`initial_refinement_level` must be a non-negative integer")
        if isinstance(periodicity, bool):
            periodicity = (periodicity,) * len(cmin)
        if not (isinstance(periodicity, (tuple, list)) and len(periodicity) == len(cmin)):
            raise ValueError("`periodicity` must be a Bool or a tuple of ndims Bools")
        self.periodicity = tuple(bool(p) for p in periodicity)
        self.ndims = len(cmin)
        self.coordinates_min = cmin
        self.coordinates_max = cmax
        self.initial_refinement_level = initial_refinement_level
        length = cmax[0] - cmin[0]
        for d in range(1, self.ndims):
            if not np.isclose(cmax[d] - cmin[d], length):
                raise ValueError("The TreeMesh domain must be a hypercube")
        self.length_level_0 = length
        self.center_level_0 = tuple((a + b) / 2 for a, b in zip(cmin, cmax))
        self.refinement_patches = []
        for patch in refinement_patches:
            if isinstance(patch, dict):
                if patch.get("type", "box") != "box":
                    raise ValueError("only refinement patches of type 'box' are supported")
                patch = (patch["coordinates_min"], patch["coordinates_max"])
            pmin, pmax = patch
            self.refinement_patches.append((tuple(float(v) for v in pmin), tuple(float(v) for v in pmax)))
        self._tree = None  # _Leaves once the tree is refined or adapted
        self.unsaved_changes = True  # mesh file out of date (TreeMesh.unsaved_changes)
        self.current_filename = ""
        self._build()
        if self.refinement_patches:
            self._build_refined()
        else:
            self.cell_lengths = np.full(self.nelements, self.cell_length)
            self.levels = np.full(self.nelements, self.initial_refinement_level, dtype=np.int32)
            self.mortar_neighbor_ids = np.zeros((0, (1 << self.ndims - 1) + 1), dtype=np.int32)
            self.mortar_large_sides = np.zeros(0, dtype=np.int32)
            self.mortar_orientations = np.zeros(0, dtype=np.int32)
            self.nmortars = 0

    def _build(self):
        D, L = self.ndims, self.initial_refinement_level
        centers, icoords = [], []

        def rec(level, center, length, ic):
            if level >= L:
                centers.append(center)
                icoords.append(ic)
                return
            child_length = length / 2
            for child in range(1 << D):
                cc, ci = [], []
                for d in range(D):
                    sgn = 1 if (child >> d) & 1 else -1
                    cc.append(center[d] + sgn * child_length / 2)
                    ci.append(2 * ic[d] + (1 if sgn > 0 else 0))
                rec(level + 1, tuple(cc), child_length, tuple(ci))

        rec(0, self.center_level_0, self.length_level_0, (0,) * D)
        self.nelements = len(centers)
        self.cell_centers = np.array(centers, dtype=np.float64)  # [E, D]
        self._icoords = icoords
        self._uniform_centers = self.cell_centers.copy()
        self.cell_length = self.length_level_0 / (1 << L)
        nside = 1 << L
        ic = np.array(icoords, dtype=np.int64)
        lin = np.zeros(self.nelements, dtype=np.int64)
        s = 1
        for d in range(D):
            lin += ic[:, d] * s
            s *= nside
        lookup = np.empty(nside ** D, dtype=np.int64)
        lookup[lin] = np.arange(self.nelements)
        # element neighbour per direction (0-based: 2o = -o, 2o+1 = +o), -1
        # across a non-periodic domain face
        nbr = np.empty((self.nelements, 2 * D), dtype=np.int32)
        for dirn in range(2 * D):
            o = dirn // 2
            shifted = ic.copy()
            shifted[:, o] += 1 if dirn % 2 else -1
            outside = (shifted[:, o] < 0) | (shifted[:, o] >= nside)
            shifted[:, o] %= nside
            l2 = np.zeros(self.nelements, dtype=np.int64)
            s = 1
            for d in range(D):
                l2 += shifted[:, d] * s
                s *= nside
            nbr[:, dirn] = lookup[l2]
            if not self.periodicity[o]:
                nbr[outside, dirn] = -1
        self.element_neighbors = nbr
        # interfaces: element order, positive directions (left = element)
        E = self.nelements
        has = nbr[:, 1::2] >= 0  # [E, D]: interface at the + face
        el, o = np.nonzero(has)  # row-major: element order, then orientation
        self.neighbor_ids = np.stack([el, nbr[el, 2 * o + 1]], axis=1).astype(np.int32)
        self.orientations = o.astype(np.int32)
        self.ninterfaces = len(o)
        # boundaries: direction-major, then element order
        b_el, b_dir = [], []
        for dirn in range(2 * D):
            ids = np.nonzero(nbr[:, dirn] < 0)[0]
            b_el.append(ids)
            b_dir.append(np.full(len(ids), dirn))
        b_el = np.concatenate(b_el).astype(np.int32)
        b_dir = np.concatenate(b_dir).astype(np.int32)
        self.boundary_neighbor_ids = b_el
        self.boundary_orientations = (b_dir // 2).astype(np.int32)
        self.boundary_neighbor_sides = np.where(b_dir % 2 == 1, 1, 2).astype(np.int32)
        self.nboundaries = len(b_el)

    def _build_refined(self):
        """Leaves of the uniform tree refined by the box patches (cells whose
        centre lies strictly inside), rebalanced 2:1, in depth-first order, and
        their interfaces, L2 mortars and boundaries."""
        nt = _NativeTree(self._uniform_tree(), self.ndims, self.length_level_0, self.periodicity)
        for pmin, pmax in self.refinement_patches:
            c = nt.leaves().centers
            nt.refine(np.all((np.array(pmin)[None, :] < c) & (np.array(pmax)[None, :] > c), axis=1))
        self._tree = nt.leaves()
        self._connect(nt)

    def _uniform_tree(self):
        ic = np.array(self._icoords, dtype=np.int64).reshape(-1, self.ndims)
        n = len(ic)
        return _Leaves(np.full(n, self.initial_refinement_level, np.int64), ic, self._uniform_centers.copy(),
                       np.zeros(n, np.int32), np.arange(n, dtype=np.int32))

    def _connect(self, nt=None):
        """Element geometry, interfaces, L2 mortars and boundaries of the leaves
        (init_interfaces!, init_mortars!, init_boundaries! of containers_{2,3}d.jl):
        a same-level leaf across a face is an interface neighbour; a same-level
        cell with children makes a mortar with this element as the large side;
        no same-level cell means this element is a small element of the coarser
        neighbour's mortar; outside a non-periodic domain the face is a boundary.
        nt: a native tree over the same leaves (one is made when None)."""
        D = self.ndims
        t = self._tree
        if nt is None:
            nt = _NativeTree(t, D, self.length_level_0, self.periodicity)
        c = nt.connectivity()
        self.nelements = len(t.lv)
        self.cell_centers = t.centers.copy()
        self.levels = t.lv.astype(np.int32)
        self.cell_lengths = self.length_level_0 / (1 << self.levels).astype(np.float64)
        self.element_neighbors = c["element_neighbors"]
        self.neighbor_ids = c["neighbor_ids"]
        self.orientations = c["orientations"]
        self.ninterfaces = len(self.orientations)
        self.mortar_neighbor_ids = c["mortar_neighbor_ids"]
        self.mortar_large_sides = c["mortar_large_sides"]
        self.mortar_orientations = c["mortar_orientations"]
        self.nmortars = len(self.mortar_orientations)
        self.boundary_neighbor_ids = c["boundary_neighbor_ids"]
        self.boundary_orientations = c["boundary_orientations"]
        self.boundary_neighbor_sides = c["boundary_neighbor_sides"]
        self.nboundaries = len(self.boundary_neighbor_ids)

    def adapt(self, lam, only_refine=False, only_coarsen=False):
        """One AMR adaptation of the tree (AMRCallback on a TreeMesh,
        amr.jl:244-392): leaves with lam > 0 are refined (refine!, with 2:1
        rebalancing, abstract_tree.jl:367-490), then parents whose 2^D children
        are all still leaves with lam < 0 are coarsened unless a child has a
        non-sibling face neighbour with children (coarsen!,
        abstract_tree.jl:516-600), highest level first.

        Returns (changed, kind, src): for every element of the adapted mesh
        kind 0 = copy of old element src, 1 + q = child q of old element src,
        -1 = parent of old elements src .. src + 2^D - 1 (the solution
        transfer map of trixi_hip_transfer)."""
        lam = np.asarray(lam)
        if lam.shape != (self.nelements,):
            raise ValueError("one controller value per element")
        t = self._tree if self._tree is not None else self._uniform_tree()
        nt = _NativeTree(t, self.ndims, self.length_level_0, self.periodicity)  # origins: kind 0, src = index
        coarsen_lv, coarsen_ic = t.lv[lam < 0], t.ic[lam < 0]
        changed = 0
        if not only_coarsen and np.any(lam > 0):
            changed += nt.refine(lam > 0)
        if not only_refine and len(coarsen_lv):
            changed += nt.coarsen(coarsen_lv, coarsen_ic)
        t = nt.leaves()
        if changed:
            self._tree = t
            self._connect(nt)
            self.unsaved_changes = True
        return bool(changed), t.kind, t.src

    def tree_arrays(self):
        """The tree as `save_mesh_file` writes it (mesh_io.jl:15-51): every cell in
        depth-first pre-order (the order refine! keeps, children right behind
        their parent), 1-based ids, 0 = none. Returns parent_ids [n],
        child_ids [n, 2^d], neighbor_ids [n, 2d] (same-level face neighbours,
        wrapping periodic dimensions; the periodic root is its own neighbour),
        levels [n], coordinates [n, d] (cell centres: the root's centre, then
        each child's = parent's +- child length / 2)."""
        D, per = self.ndims, self.periodicity
        t = self._tree if self._tree is not None else self._uniform_tree()
        cl = _Cells(t.lv, t.ic)
        lookup = dict(zip(cl.k.tolist(), cl.v.tolist()))

        class _V:  # scalar lookups of the cell table
            @staticmethod
            def value(lv, ic):
                k = lv
                for d in range(D):
                    k |= ic[d] << _KEY_SHIFT[d]
                return lookup.get(k)

        cells = _V
        nch = 1 << D
        order, coords = [], []
        stack = [(0, (0,) * D, self.center_level_0)]
        while stack:
            lv, ic, c = stack.pop()
            order.append((lv, ic))
            coords.append(c)
            if cells.value(lv, ic) < 0:
                child_length = self.length_level_0 / (1 << lv) / 2
                for q in reversed(range(nch)):
                    bits = [(q >> d) & 1 for d in range(D)]
                    stack.append((lv + 1, tuple(2 * ic[d] + bits[d] for d in range(D)),
                                  tuple(c[d] + (1 if bits[d] else -1) * child_length / 2 for d in range(D))))
        ids = {key: k + 1 for k, key in enumerate(order)}
        n = len(order)
        parent = np.zeros(n, np.int64)
        child = np.zeros((n, nch), np.int64)
        nbr = np.zeros((n, 2 * D), np.int64)
        levels = np.array([lv for lv, _ in order], np.int64)
        for k, (lv, ic) in enumerate(order):
            if lv > 0:
                parent[k] = ids[(lv - 1, tuple(v >> 1 for v in ic))]
            if cells.value(lv, ic) < 0:
                for q in range(nch):
                    child[k, q] = ids[(lv + 1, tuple(2 * ic[d] + ((q >> d) & 1) for d in range(D)))]
            for dirn in range(2 * D):
                o, nside = dirn // 2, 1 << lv
                kk = ic[o] + (1 if dirn % 2 else -1)
                if (kk < 0 or kk >= nside) and not per[o]:
                    continue
                nb = list(ic)
                nb[o] = kk % nside
                nb = tuple(nb)
                if (lv, nb) in ids:
                    nbr[k, dirn] = ids[(lv, nb)]
        return parent, child, nbr, levels, np.array(coords, np.float64).reshape(n, D)

    @classmethod
    def from_tree(cls, center_level_0, length_level_0, periodicity, levels, coordinates, child_ids):
        """The TreeMesh of a tree as `load_mesh!` reads it (mesh_io.jl:603-628):
        its leaves are the cells without children, in id order."""
        D = len(center_level_0)
        leaf = np.all(np.asarray(child_ids) == 0, axis=1)
        lv = np.asarray(levels)[leaf].astype(np.int64)
        xc = np.asarray(coordinates, np.float64)[leaf].reshape(-1, D)
        L0 = float(length_level_0)
        cmin = tuple(float(c) - L0 / 2 for c in center_level_0)
        cmax = tuple(float(c) + L0 / 2 for c in center_level_0)
        mesh = cls(cmin, cmax, initial_refinement_level=int(lv.min()), periodicity=tuple(bool(p) for p in periodicity))
        h = L0 / (1 << lv).astype(np.float64)
        ic = np.rint((xc - np.array(cmin)[None, :]) / h[:, None] - 0.5).astype(np.int64)
        n = len(lv)
        mesh._tree = _Leaves(lv, ic, xc.copy(), np.zeros(n, np.int32), np.arange(n, dtype=np.int32))
        mesh._connect()
        mesh.unsaved_changes = False
        return mesh

    def ncells(self):
        return self.nelements

    def __repr__(self):
        return (f"TreeMesh{{{self.ndims}}} with {self.nelements} elements on "
                f"{self.coordinates_min}..{self.coordinates_max}, periodicity={self.periodicity}")


class _Leaves:
    """The leaves of the tree in element (depth-first) order: level, integer
    coordinates at that level, centre, and the origin of each leaf in the
    current adaptation (kind 0 = old leaf src, 1 + q = child q of old leaf src,
    -1 = parent of old leaves src .. src + 2^D - 1)."""

    def __init__(self, lv, ic, centers, kind, src):
        self.lv, self.ic, self.centers, self.kind, self.src = lv, ic, centers, kind, src



_KEY_SHIFT = (6, 25, 44)  # level in bits 0-5, coordinate d in 19 bits from _KEY_SHIFT[d]


def _keys(lv, ic):
    k = np.asarray(lv, dtype=np.int64).copy()
    for d in range(ic.shape[1]):
        k |= np.asarray(ic[:, d], dtype=np.int64) << _KEY_SHIFT[d]
    return k


class _Cells:
    """Every cell of the tree by key (level, integer coordinates), sorted: a
    leaf maps to its element index, an inner cell (one with children) to -1
    (the mesh file's tree walk, tree_arrays)."""

    def __init__(self, lv, ic):
        keys, vals = [_keys(lv, ic)], [np.arange(len(lv), dtype=np.int64)]
        for level in range(int(lv.max()) if len(lv) else 0):
            m = lv > level
            if m.any():
                a = np.unique(_keys(np.full(int(m.sum()), level), ic[m] >> (lv[m] - level)[:, None]))
                keys.append(a)
                vals.append(np.full(len(a), -1, dtype=np.int64))
        k = np.concatenate(keys)
        v = np.concatenate(vals)
        o = np.argsort(k, kind="stable")
        self.k, self.v = k[o], v[o]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class _NativeTree:
    """A native tree (trixi_hip_tree) over the leaves of a _Leaves: refine
    with rebalancing, coarsen, leaves with their origin, connectivity."""

    def __init__(self, t, D, length0, per):
        self._lib = _lib.load()
        self.D = D
        n = len(t.lv)
        lv = np.ascontiguousarray(t.lv, dtype=np.int32)
        ic = np.ascontiguousarray(np.asarray(t.ic, dtype=np.int64).reshape(n, D))
        c = np.ascontiguousarray(np.asarray(t.centers, dtype=np.float64).reshape(n, D))
        p = (ctypes.c_int * D)(*[int(bool(x)) for x in per])
        self.h = self._lib.trixi_hip_tree_create(D, p, float(length0), n, _ptr(lv), _ptr(ic), _ptr(c))
        if not self.h:
            raise ValueError("invalid tree leaves (levels 0..19, coordinates inside their level)")

    def __del__(self):
        h, self.h = getattr(self, "h", None), None
        if h:
            self._lib.trixi_hip_tree_destroy(h)

    def _check(self, rc, what):
        if rc != _lib.OK:
            raise ValueError(f"{what}: {self._lib.trixi_hip_tree_last_error(self.h).decode()}")

    def refine(self, marked):
        m = np.ascontiguousarray(np.asarray(marked, dtype=bool).astype(np.uint8))
        c = ctypes.c_int64()
        self._check(self._lib.trixi_hip_tree_refine(self.h, _ptr(m), ctypes.byref(c)), "refine")
        return int(c.value)

    def coarsen(self, lv, ic):
        lv = np.ascontiguousarray(lv, dtype=np.int32)
        ic = np.ascontiguousarray(np.asarray(ic, dtype=np.int64).reshape(len(lv), self.D))
        c = ctypes.c_int64()
        self._check(self._lib.trixi_hip_tree_coarsen(self.h, len(lv), _ptr(lv), _ptr(ic), ctypes.byref(c)),
                    "coarsen")
        return int(c.value)

    def sizes(self):
        s = np.zeros(4, dtype=np.int64)
        self._check(self._lib.trixi_hip_tree_sizes(self.h, _ptr(s)), "tree sizes")
        return [int(v) for v in s]

    def leaves(self):
        n = self.sizes()[0]
        lv = np.empty(n, np.int32)
        ic = np.empty((n, self.D), np.int64)
        c = np.empty((n, self.D), np.float64)
        kind = np.empty(n, np.int32)
        src = np.empty(n, np.int32)
        self._check(self._lib.trixi_hip_tree_leaves(self.h, _ptr(lv), _ptr(ic), _ptr(c), _ptr(kind), _ptr(src)),
                    "tree leaves")
        return _Leaves(lv.astype(np.int64), ic, c, kind, src)

    def connectivity(self):
        n, ni, nm, nb = self.sizes()
        D, ns = self.D, 1 << (self.D - 1)
        out = dict(element_neighbors=np.empty((n, 2 * D), np.int32), neighbor_ids=np.empty((ni, 2), np.int32),
                   orientations=np.empty(ni, np.int32), mortar_neighbor_ids=np.empty((nm, ns + 1), np.int32),
                   mortar_large_sides=np.empty(nm, np.int32), mortar_orientations=np.empty(nm, np.int32),
                   boundary_neighbor_ids=np.empty(nb, np.int32), boundary_orientations=np.empty(nb, np.int32),
                   boundary_neighbor_sides=np.empty(nb, np.int32))
        order = ("element_neighbors", "neighbor_ids", "orientations", "mortar_neighbor_ids", "mortar_large_sides",
                 "mortar_orientations", "boundary_neighbor_ids", "boundary_orientations", "boundary_neighbor_sides")
        self._check(self._lib.trixi_hip_tree_connectivity(self.h, *[_ptr(out[k]) for k in order]), "connectivity")
        return out


def element_geometry(mesh, basis):
    """inverse_jacobian[E] and node_coordinates[E, NN, D] as init_elements! computes them."""
    D = mesh.ndims
    n = basis.nnodes()
    reference_length = float(np.sum(basis.weights))
    reference_offset = (basis.nodes[0] + basis.nodes[-1]) / 2
    jacobian = mesh.cell_lengths / reference_length  # [E]
    inv_jac = 1.0 / jacobian
    rel = jacobian[:, None] * (basis.nodes - reference_offset)[None, :]  # [E, n]
    NN = n ** D
    idx = np.indices((n,) * D).reshape(D, -1)[::-1]  # idx[d, node], node = i + n j (+ n^2 k)
    coords = np.empty((mesh.nelements, NN, D))
    for d in range(D):
        coords[:, :, d] = mesh.cell_centers[:, d][:, None] + rel[:, idx[d]]
    return inv_jac, coords
