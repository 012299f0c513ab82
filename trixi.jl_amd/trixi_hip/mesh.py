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
        self._leaves = None  # (level, integer coordinates, centre) per leaf once the tree is refined
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
        leaves = self._uniform_leaves()
        for pmin, pmax in self.refinement_patches:
            marked = {i for i, (_, _, c) in enumerate(leaves)
                      if all(pmin[d] < c[d] and pmax[d] > c[d] for d in range(D))}
            leaves, _, _ = _refine_balanced(leaves, [(0, i) for i in range(len(leaves))], marked, D,
                                            self.length_level_0, self.periodicity)
        self._leaves = leaves
        self._connect()

    def _uniform_leaves(self):
        L = self.initial_refinement_level
        return [(L, tuple(int(v) for v in ic), tuple(float(v) for v in c))
                for ic, c in zip(self._icoords, self.cell_centers)]

    def _connect(self):
        """Element geometry, interfaces, L2 mortars and boundaries of the leaves."""
        D, per = self.ndims, self.periodicity
        leaves = self._leaves
        cells = _tree_cells(leaves)
        E = len(leaves)
        self.nelements = E
        self.cell_centers = np.array([c for _, _, c in leaves], dtype=np.float64).reshape(E, D)
        self.levels = np.array([lv for lv, _, _ in leaves], dtype=np.int32)
        self.cell_lengths = self.length_level_0 / (1 << self.levels).astype(np.float64)
        ns = 1 << (D - 1)
        nbr = np.full((E, 2 * D), -1, dtype=np.int32)
        boundary = np.zeros((E, 2 * D), dtype=bool)
        m_ids, m_side, m_ori = [], [], []
        for e, (lv, ic, _) in enumerate(leaves):
            for dirn in range(2 * D):
                nb = _neighbour(lv, ic, dirn, per)
                if nb is None:
                    boundary[e, dirn] = True
                    continue
                f = cells.get((lv, nb))
                if f is None:
                    continue  # small side of a coarser neighbour's mortar
                if f >= 0:
                    nbr[e, dirn] = f
                    continue
                o, plus = dirn // 2, dirn % 2
                free = [d for d in range(D) if d != o]
                ids = []
                for q in range(ns):
                    ch = [2 * v for v in nb]
                    ch[o] += 0 if plus else 1
                    ch[free[0]] += q & 1
                    if D == 3:
                        ch[free[1]] += q >> 1
                    f2 = cells.get((lv + 1, tuple(ch)))
                    if f2 is None or f2 < 0:
                        raise ValueError("the tree is not 2:1 balanced")
                    ids.append(f2)
                m_ids.append(ids + [e])
                m_side.append(1 if plus else 2)
                m_ori.append(o)
        self.element_neighbors = nbr
        has = nbr[:, 1::2] >= 0
        el, o = np.nonzero(has)
        self.neighbor_ids = np.stack([el, nbr[el, 2 * o + 1]], axis=1).astype(np.int32).reshape(-1, 2)
        self.orientations = o.astype(np.int32)
        self.ninterfaces = len(o)
        self.mortar_neighbor_ids = np.array(m_ids, dtype=np.int32).reshape(-1, ns + 1)
        self.mortar_large_sides = np.array(m_side, dtype=np.int32)
        self.mortar_orientations = np.array(m_ori, dtype=np.int32)
        self.nmortars = len(m_side)
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
        leaves = self._leaves if self._leaves is not None else self._uniform_leaves()
        orig = [(0, i) for i in range(len(leaves))]
        to_refine = set(np.nonzero(lam > 0)[0].tolist())
        to_coarsen = [(leaves[i][0], leaves[i][1]) for i in np.nonzero(lam < 0)[0].tolist()]
        changed = 0
        if not only_coarsen and to_refine:
            leaves, orig, c = _refine_balanced(leaves, orig, to_refine, D, self.length_level_0, self.periodicity)
            changed += c
        if not only_refine and to_coarsen:
            leaves, orig, c = _coarsen(leaves, orig, to_coarsen, D, self.length_level_0, self.periodicity)
            changed += c
        if changed:
            self._leaves = leaves
            self._connect()
            self.unsaved_changes = True
        kind = np.array([o[0] for o in orig], dtype=np.int32)
        src = np.array([o[1] for o in orig], dtype=np.int32)
        return bool(changed), kind, src

    def tree_arrays(self):
        """The tree as `save_mesh_file` writes it (mesh_io.jl:15-51): every cell in
        depth-first pre-order (the order refine! keeps, children right behind
        their parent), 1-based ids, 0 = none. Returns parent_ids [n],
        child_ids [n, 2^d], neighbor_ids [n, 2d] (same-level face neighbours,
        wrapping periodic dimensions; the periodic root is its own neighbour),
        levels [n], coordinates [n, d] (cell centres: the root's centre, then
        each child's = parent's +- child length / 2)."""
        D, per = self.ndims, self.periodicity
        leaves = self._leaves if self._leaves is not None else self._uniform_leaves()
        cells = _tree_cells(leaves)
        nch = 1 << D
        order, coords = [], []
        stack = [(0, (0,) * D, self.center_level_0)]
        while stack:
            lv, ic, c = stack.pop()
            order.append((lv, ic))
            coords.append(c)
            if cells[(lv, ic)] < 0:
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
            if cells[(lv, ic)] < 0:
                for q in range(nch):
                    child[k, q] = ids[(lv + 1, tuple(2 * ic[d] + ((q >> d) & 1) for d in range(D)))]
            for dirn in range(2 * D):
                nb = _neighbour(lv, ic, dirn, per)
                if nb is not None and (lv, nb) in ids:
                    nbr[k, dirn] = ids[(lv, nb)]
        return parent, child, nbr, levels, np.array(coords, np.float64).reshape(n, D)

    @classmethod
    def from_tree(cls, center_level_0, length_level_0, periodicity, levels, coordinates, child_ids):
        """The TreeMesh of a tree as `load_mesh!` reads it (mesh_io.jl:603-628):
        its leaves are the cells without children, in id order."""
        D = len(center_level_0)
        leaf = np.all(np.asarray(child_ids) == 0, axis=1)
        lv = np.asarray(levels)[leaf].astype(int)
        xc = np.asarray(coordinates, np.float64)[leaf]
        L0 = float(length_level_0)
        cmin = tuple(float(c) - L0 / 2 for c in center_level_0)
        cmax = tuple(float(c) + L0 / 2 for c in center_level_0)
        mesh = cls(cmin, cmax, initial_refinement_level=int(lv.min()), periodicity=tuple(bool(p) for p in periodicity))
        leaves = []
        for level, c in zip(lv.tolist(), xc):
            h = L0 / (1 << level)
            ic = tuple(int(round((c[d] - cmin[d]) / h - 0.5)) for d in range(D))
            leaves.append((level, ic, tuple(float(v) for v in c)))
        mesh._leaves = leaves
        mesh._connect()
        mesh.unsaved_changes = False
        return mesh

    def ncells(self):
        return self.nelements

    def __repr__(self):
        return (f"TreeMesh{{{self.ndims}}} with {self.nelements} elements on "
                f"{self.coordinates_min}..{self.coordinates_max}, periodicity={self.periodicity}")


def _tree_cells(leaves):
    """Every cell of the tree by (level, integer coordinates): a leaf maps to its
    element index, an inner cell (one with children) to -1."""
    cells = {}
    for i, (lv, ic, _) in enumerate(leaves):
        cells[(lv, ic)] = i
        a = ic
        for l2 in range(lv - 1, -1, -1):
            a = tuple(v >> 1 for v in a)
            if (l2, a) in cells:
                break
            cells[(l2, a)] = -1
    return cells


def _neighbour(lv, ic, dirn, per):
    """Same-level cell coordinates across face dirn (2o = -o, 2o+1 = +o),
    wrapping periodic dimensions; None outside a non-periodic domain."""
    o, nside = dirn // 2, 1 << lv
    k = ic[o] + (1 if dirn % 2 else -1)
    if (k < 0 or k >= nside) and not per[o]:
        return None
    nb = list(ic)
    nb[o] = k % nside
    return tuple(nb)


def _refine(leaves, orig, marked, D, length0):
    """refine_unbalanced! (abstract_tree.jl:317-353): each marked leaf is
    replaced in place by its children (bit d of the child index = upper half in d)."""
    out, oout = [], []
    for i, leaf in enumerate(leaves):
        if i not in marked:
            out.append(leaf)
            oout.append(orig[i])
            continue
        lv, ic, c = leaf
        child_length = length0 / (1 << lv) / 2
        for child in range(1 << D):
            bits = [(child >> d) & 1 for d in range(D)]
            out.append((lv + 1, tuple(2 * ic[d] + bits[d] for d in range(D)),
                        tuple(c[d] + (1 if bits[d] else -1) * child_length / 2 for d in range(D))))
            oout.append((1 + child, orig[i][1]))
    return out, oout


def _refine_balanced(leaves, orig, marked, D, length0, per):
    """refine! + rebalance! (abstract_tree.jl:367-400, 453-490): after each
    round, the coarse neighbour (the parent's face neighbour) of a refined cell
    without a same-level face neighbour is refined, until nothing changes."""
    count = 0
    while marked:
        refined = [leaves[i] for i in sorted(marked)]
        count += len(refined)
        leaves, orig = _refine(leaves, orig, marked, D, length0)
        cells = _tree_cells(leaves)
        marked = set()
        for lv, ic, _ in refined:
            for dirn in range(2 * D):
                nb = _neighbour(lv, ic, dirn, per)
                if nb is not None and (lv, nb) in cells:
                    continue
                if lv == 0:
                    continue
                pnb = _neighbour(lv - 1, tuple(v >> 1 for v in ic), dirn, per)
                if pnb is None:
                    continue
                f = cells.get((lv - 1, pnb))
                if f is not None and f >= 0:
                    marked.add(f)
    return leaves, orig, count


def _coarsen(leaves, orig, keys, D, length0, per):
    """Coarsening of the AMR callback (amr.jl:304-370) with the balance guard
    of coarsen! (abstract_tree.jl:516-600)."""
    nch = 1 << D
    cells = _tree_cells(leaves)
    count, parents = {}, []
    for lv, ic in keys:
        if lv == 0:
            continue
        f = cells.get((lv, ic))
        if f is None or f < 0:
            continue  # refined by rebalancing: no longer a leaf
        pk = (lv - 1, tuple(v >> 1 for v in ic))
        count[pk] = count.get(pk, 0) + 1
        if count[pk] == nch:
            parents.append(pk)
    parents.sort(key=lambda p: -p[0])
    coarsened = set()
    for lv, pic in parents:
        children = [tuple(2 * pic[d] + ((q >> d) & 1) for d in range(D)) for q in range(nch)]
        skip = False
        for q, ch in enumerate(children):
            for dirn in range(2 * D):
                bit = (q >> (dirn // 2)) & 1
                if bit == (1 if dirn % 2 == 0 else 0):
                    continue  # the neighbour is a sibling
                nb = _neighbour(lv + 1, ch, dirn, per)
                if nb is None:
                    continue
                f = cells.get((lv + 1, nb))
                if f is not None and f < 0:
                    skip = True
                    break
            if skip:
                break
        if skip:
            continue
        for ch in children:
            del cells[(lv + 1, ch)]
        cells[(lv, pic)] = 0
        coarsened.add((lv, pic))
    if not coarsened:
        return leaves, orig, 0
    out, oout = [], []
    i = 0
    while i < len(leaves):
        lv, ic, c = leaves[i]
        if lv > 0:
            pic = tuple(v >> 1 for v in ic)
            if (lv - 1, pic) in coarsened:
                h = length0 / (1 << lv)
                out.append((lv - 1, pic, tuple(c[d] + h / 2 for d in range(D))))
                oout.append((-1, orig[i][1]))
                i += nch
                continue
        out.append(leaves[i])
        oout.append(orig[i])
        i += 1
    return out, oout, len(coarsened)


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
