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
"""
import numpy as np


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
        D = self.ndims
        t = self._uniform_tree()
        for pmin, pmax in self.refinement_patches:
            c = t.centers
            marked = np.all((np.array(pmin)[None, :] < c) & (np.array(pmax)[None, :] > c), axis=1)
            t, _ = _refine_balanced(t, marked, D, self.length_level_0, self.periodicity)
        self._tree = t
        self._connect()

    def _uniform_tree(self):
        ic = np.array(self._icoords, dtype=np.int64).reshape(-1, self.ndims)
        n = len(ic)
        return _Leaves(np.full(n, self.initial_refinement_level, np.int64), ic, self._uniform_centers.copy(),
                       np.zeros(n, np.int32), np.arange(n, dtype=np.int32))

    def _connect(self):
        """Element geometry, interfaces, L2 mortars and boundaries of the leaves
        (init_interfaces!, init_mortars!, init_boundaries! of containers_{2,3}d.jl):
        a same-level leaf across a face is an interface neighbour; a same-level
        cell with children makes a mortar with this element as the large side;
        no same-level cell means this element is a small element of the coarser
        neighbour's mortar; outside a non-periodic domain the face is a boundary."""
        D, per = self.ndims, self.periodicity
        t = self._tree
        E = len(t.lv)
        cells = _Cells(t.lv, t.ic)
        self.nelements = E
        self.cell_centers = t.centers.copy()
        self.levels = t.lv.astype(np.int32)
        self.cell_lengths = self.length_level_0 / (1 << self.levels).astype(np.float64)
        ns = 1 << (D - 1)
        nbr = np.full((E, 2 * D), -1, dtype=np.int32)
        boundary = np.zeros((E, 2 * D), dtype=bool)
        inner_nb = np.zeros((E, 2 * D), dtype=bool)
        nb_ic = np.zeros((2 * D, E, D), dtype=np.int64)
        for dirn in range(2 * D):
            nb, valid = _shift(t.lv, t.ic, dirn, per)
            boundary[:, dirn] = ~valid
            found, val = cells.find(t.lv, nb)
            leaf = valid & found & (val >= 0)
            nbr[leaf, dirn] = val[leaf]
            inner_nb[:, dirn] = valid & found & (val < 0)
            nb_ic[dirn] = nb
        # mortars in element order, then direction order
        m_el, m_dir = np.nonzero(inner_nb)
        M = len(m_el)
        m_ids = np.zeros((M, ns + 1), dtype=np.int32)
        m_ids[:, ns] = m_el
        o = m_dir // 2
        plus = m_dir % 2
        nbm = nb_ic[m_dir, m_el]  # [M, D] same-level neighbour cell (has children)
        lvm = t.lv[m_el] + 1
        free = np.array([[d for d in range(D) if d != oo] for oo in range(D)], dtype=np.int64)[o]
        rows = np.arange(M)
        for q in range(ns):
            ch = 2 * nbm
            ch[rows, o] += 1 - plus
            ch[rows, free[:, 0]] += q & 1
            if D == 3:
                ch[rows, free[:, 1]] += q >> 1
            found, val = cells.find(lvm, ch)
            if not np.all(found & (val >= 0)):
                raise ValueError("the tree is not 2:1 balanced")
            m_ids[:, q] = val
        self.element_neighbors = nbr
        has = nbr[:, 1::2] >= 0
        el, oi = np.nonzero(has)
        self.neighbor_ids = np.stack([el, nbr[el, 2 * oi + 1]], axis=1).astype(np.int32).reshape(-1, 2)
        self.orientations = oi.astype(np.int32)
        self.ninterfaces = len(oi)
        self.mortar_neighbor_ids = m_ids
        self.mortar_large_sides = np.where(plus == 1, 1, 2).astype(np.int32)
        self.mortar_orientations = o.astype(np.int32)
        self.nmortars = M
        b_el, b_dir = [], []
        for dirn in range(2 * D):
            ids = np.nonzero(boundary[:, dirn])[0]
            b_el.append(ids)
            b_dir.append(np.full(len(ids), dirn))
        b_el = np.concatenate(b_el).astype(np.int32)
        b_dir = np.concatenate(b_dir).astype(np.int32)
        self.boundary_neighbor_ids = b_el
        self.boundary_orientations = (b_dir // 2).astype(np.int32)
        self.boundary_neighbor_sides = np.where(b_dir % 2 == 1, 1, 2).astype(np.int32)
        self.nboundaries = len(b_el)

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
        D = self.ndims
        t = self._tree if self._tree is not None else self._uniform_tree()
        t = _Leaves(t.lv, t.ic, t.centers, np.zeros(len(t.lv), np.int32), np.arange(len(t.lv), dtype=np.int32))
        coarsen_lv, coarsen_ic = t.lv[lam < 0], t.ic[lam < 0]
        changed = 0
        if not only_coarsen and np.any(lam > 0):
            t, c = _refine_balanced(t, lam > 0, D, self.length_level_0, self.periodicity)
            changed += c
        if not only_refine and len(coarsen_lv):
            t, c = _coarsen(t, coarsen_lv, coarsen_ic, D, self.length_level_0, self.periodicity)
            changed += c
        if changed:
            self._tree = t
            self._connect()
            self.unsaved_changes = True
        return bool(changed), t.kind.astype(np.int32), t.src.astype(np.int32)

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

    def take(self, idx):
        return _Leaves(self.lv[idx], self.ic[idx], self.centers[idx], self.kind[idx], self.src[idx])


_KEY_SHIFT = (6, 25, 44)  # level in bits 0-5, coordinate d in 19 bits from _KEY_SHIFT[d]


def _keys(lv, ic):
    k = np.asarray(lv, dtype=np.int64).copy()
    for d in range(ic.shape[1]):
        k |= np.asarray(ic[:, d], dtype=np.int64) << _KEY_SHIFT[d]
    return k


class _Cells:
    """Every cell of the tree by key (level, integer coordinates): a leaf maps
    to its element index, an inner cell (one with children) to -1."""

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

    def find(self, lv, ic):
        """(found, value) for every query cell; value -2 where absent."""
        q = _keys(lv, ic)
        if len(self.k) == 0:
            return np.zeros(len(q), bool), np.full(len(q), -2, np.int64)
        pos = np.minimum(np.searchsorted(self.k, q), len(self.k) - 1)
        found = self.k[pos] == q
        return found, np.where(found, self.v[pos], -2)

    def value(self, lv, ic):
        f, v = self.find(np.array([lv]), np.array([ic], dtype=np.int64).reshape(1, -1))
        return int(v[0]) if f[0] else None


def _shift(lv, ic, dirn, per):
    """Same-level cell coordinates across face dirn (2o = -o, 2o+1 = +o),
    wrapping periodic dimensions; valid is False outside a non-periodic domain."""
    o = dirn // 2
    nside = np.left_shift(1, lv)
    k = ic[:, o] + (1 if dirn % 2 else -1)
    outside = (k < 0) | (k >= nside)
    nb = ic.copy()
    nb[:, o] = np.mod(k, nside)
    valid = np.ones(len(lv), bool) if per[o] else ~outside
    return nb, valid


def _refine(t, marked, D, length0):
    """refine_unbalanced! (abstract_tree.jl:317-353): each marked leaf is
    replaced in place by its children (bit d of the child index = upper half in d)."""
    nch = 1 << D
    E = len(t.lv)
    counts = np.where(marked, nch, 1)
    idx = np.repeat(np.arange(E), counts)
    start = np.cumsum(counts) - counts
    q = np.arange(len(idx)) - start[idx]
    child = marked[idx]
    bits = (q[:, None] >> np.arange(D)[None, :]) & 1
    lv = t.lv[idx] + child
    ic = np.where(child[:, None], 2 * t.ic[idx] + bits, t.ic[idx])
    child_length = length0 / (1 << t.lv[idx]).astype(np.float64) / 2
    c = np.where(child[:, None], t.centers[idx] + np.where(bits == 1, 1.0, -1.0) * (child_length / 2)[:, None],
                 t.centers[idx])
    kind = np.where(child, 1 + q, t.kind[idx]).astype(np.int32)
    return _Leaves(lv, ic, c, kind, t.src[idx].copy())


def _refine_balanced(t, marked, D, length0, per):
    """refine! + rebalance! (abstract_tree.jl:367-400, 453-490): after each
    round, the coarse neighbour (the parent's face neighbour) of a refined cell
    without a same-level face neighbour is refined, until nothing changes."""
    count = 0
    marked = np.asarray(marked, bool)
    while marked.any():
        r_lv, r_ic = t.lv[marked], t.ic[marked]
        count += int(marked.sum())
        t = _refine(t, marked, D, length0)
        cells = _Cells(t.lv, t.ic)
        marked = np.zeros(len(t.lv), bool)
        for dirn in range(2 * D):
            nb, valid = _shift(r_lv, r_ic, dirn, per)
            found, _ = cells.find(r_lv, nb)
            cand = ~(valid & found) & (r_lv > 0)
            if not cand.any():
                continue
            plv, pic = r_lv[cand] - 1, r_ic[cand] >> 1
            pnb, pvalid = _shift(plv, pic, dirn, per)
            f2, v2 = cells.find(plv, pnb)
            sel = pvalid & f2 & (v2 >= 0)
            marked[v2[sel]] = True
    return t, count


def _coarsen(t, keys_lv, keys_ic, D, length0, per):
    """Coarsening of the AMR callback (amr.jl:304-370) with the balance guard
    of coarsen! (abstract_tree.jl:516-600), highest level first (within one
    level the decisions do not depend on each other)."""
    nch = 1 << D
    cells = _Cells(t.lv, t.ic)
    found, val = cells.find(keys_lv, keys_ic)
    ok = (keys_lv > 0) & found & (val >= 0)  # refined by rebalancing: no longer a leaf
    plv, pic = keys_lv[ok] - 1, keys_ic[ok] >> 1
    pk = _keys(plv, pic)
    uniq, first, cnt = np.unique(pk, return_index=True, return_counts=True)
    full = first[cnt == nch]
    plv, pic = plv[full], pic[full]
    total = 0
    for level in sorted(set(plv.tolist()), reverse=True):
        at = plv == level
        P_lv, P_ic = plv[at], pic[at]
        skip = np.zeros(len(P_lv), bool)
        for q in range(nch):
            bits = np.array([(q >> d) & 1 for d in range(D)], dtype=np.int64)
            ch = 2 * P_ic + bits[None, :]
            for dirn in range(2 * D):
                bit = (q >> (dirn // 2)) & 1
                if bit == (1 if dirn % 2 == 0 else 0):
                    continue  # the neighbour is a sibling
                nb, valid = _shift(P_lv + 1, ch, dirn, per)
                f, v = cells.find(P_lv + 1, nb)
                skip |= valid & f & (v < 0)
        done = _keys(P_lv[~skip], P_ic[~skip])
        if len(done) == 0:
            continue
        total += len(done)
        # collapse the children of the coarsened parents (consecutive, child 0 first)
        par = np.where(t.lv > 0, _keys(t.lv - 1, t.ic >> 1), -1)
        inpar = (t.lv > 0) & np.isin(par, done)
        first_child = inpar & np.all((t.ic & 1) == 0, axis=1)
        keep = ~inpar | first_child
        h = length0 / (1 << t.lv).astype(np.float64)
        lv = np.where(first_child, t.lv - 1, t.lv)
        ic = np.where(first_child[:, None], t.ic >> 1, t.ic)
        c = np.where(first_child[:, None], t.centers + (h / 2)[:, None], t.centers)
        kind = np.where(first_child, -1, t.kind).astype(np.int32)
        t = _Leaves(lv, ic, c, kind, t.src).take(keep)
        cells = _Cells(t.lv, t.ic)
    return t, total


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
