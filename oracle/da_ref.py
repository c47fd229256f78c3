"""Brute-force restatement of the reference's octree and DA construction for small meshes.

TEST INFRASTRUCTURE ONLY (used by tests/ as the checker of the C++ host builder).
Pure Python/numpy, O(n^2) where convenient; meant for a few thousand octants.

Octants are (x, y, z, level) tuples in the coordinates of an octree of max depth D (an
octant of level l has side 2^(D-l)).  Paths cite the reference tree.
"""
import numpy as np


def morton_key(x, y, z):
    """Interleaved key, z most significant within a bit triple (TreeNode.txx:160-180)."""
    k = 0
    for b in range(21):
        k |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
    return k


def sort_octs(octs):
    return sorted(octs, key=lambda o: (morton_key(o[0], o[1], o[2]), o[3]))


def side(D, l):
    return 1 << (D - l)


def child_number(o, D):
    """Morton child number 4k + 2j + i (include/oda/oda.txx:1087-1109)."""
    x, y, z, l = o
    if l == 0:
        return 0
    s = side(D, l)
    return (((z // s) & 1) << 2) | (((y // s) & 1) << 1) | ((x // s) & 1)


def parent(o, D):
    x, y, z, l = o
    s = side(D, l - 1)
    return (x // s * s, y // s * s, z // s * s, l - 1)


# ------------------------------------------------------------------ construction
def points_to_octree(cells, D, max_pts=1):
    """Recursive subdivision of the root until every leaf holds <= max_pts point cells
    (points2OctreeSeq + p2oLocal, src/oct/Construct.C:884-1010).  cells: (x,y,z) at depth D."""
    cells = sorted(set(map(tuple, cells)), key=lambda c: morton_key(*c))
    out = []

    def rec(o, pts):
        if len(pts) <= max_pts or o[3] == D:
            out.append(o)
            return
        s = side(D, o[3] + 1)
        for c in range(8):
            ch = (o[0] + (c & 1) * s, o[1] + ((c >> 1) & 1) * s, o[2] + ((c >> 2) & 1) * s, o[3] + 1)
            sub = [p for p in pts if ch[0] <= p[0] < ch[0] + s and ch[1] <= p[1] < ch[1] + s
                   and ch[2] <= p[2] < ch[2] + s]
            rec(ch, sub)

    rec((0, 0, 0, 0), cells)
    return sort_octs(out)


def _touch(a, b, D):
    """Closed boxes of two octants intersect (face, edge or corner contact or overlap)."""
    sa, sb = side(D, a[3]), side(D, b[3])
    for d in range(3):
        if a[d] + sa < b[d] or b[d] + sb < a[d]:
            return False
    return True


def balance(octs, D):
    """Minimal 2:1 balanced refinement (corners included) by repeated forced splitting:
    while a leaf touches a leaf more than one level finer, split it."""
    leaves = set(octs)
    changed = True
    while changed:
        changed = False
        lst = list(leaves)
        for a in lst:
            if a not in leaves:
                continue
            for b in lst:
                if b in leaves and b[3] > a[3] + 1 and _touch(a, b, D):
                    leaves.discard(a)
                    s = side(D, a[3] + 1)
                    for c in range(8):
                        leaves.add((a[0] + (c & 1) * s, a[1] + ((c >> 1) & 1) * s,
                                    a[2] + ((c >> 2) & 1) * s, a[3] + 1))
                    changed = True
                    break
    return sort_octs(leaves)


def coarsen(octs, D):
    """ot::coarsenOctree (src/oct/Coarsen.C:53-106): between consecutive child-0 octants,
    a run of at least 8 octants starts with a complete family, which becomes its parent."""
    if len(octs) < 8:
        return list(octs)
    out, prev, n = [], 0, len(octs)
    while prev < n:
        cur = prev + 1
        while cur < n and (octs[cur][3] == 0 or child_number(octs[cur], D) != 0):
            cur += 1
        if cur >= prev + 8:
            out.append(parent(octs[prev], D))
            out.extend(octs[prev + 8:cur])
        else:
            out.extend(octs[prev:cur])
        prev = cur
    return out


# ------------------------------------------------------------------ DA
class DA:
    """ot::DA on one process (src/oda/private/odaFactory.C:143-330,
    src/oda/private/odaBuildNlist.C)."""

    def __init__(self, octs, D):
        self.D = D
        Dd = D + 1
        self.Dda = Dd
        lim = 1 << D
        elems = [(x, y, z, l + 1) for (x, y, z, l) in octs]      # embed one level deeper
        pseudo = set()
        for (x, y, z, l) in octs:                                # addBoundaryNodesType1
            s = side(D, l)
            px, py, pz = x + s == lim, y + s == lim, z + s == lim
            L = l + 1
            if px:
                pseudo.add((lim, y, z, L))
            if py:
                pseudo.add((x, lim, z, L))
            if pz:
                pseudo.add((x, y, lim, L))
            if px and py:
                pseudo.add((lim, lim, z, L))
            if px and pz:
                pseudo.add((lim, y, lim, L))
            if py and pz:
                pseudo.add((x, lim, lim, L))
            if px and py and pz:
                pseudo.add((lim, lim, lim, L))
        self.elems = elems
        self.octs = elems + sort_octs(pseudo)
        self.E = len(elems)
        allset = set(self.octs)
        # flagNodesType3 (src/oct/OctFunctions.C:1501-1640): keys of the parent-level
        # neighbours across the edge/face the anchor lies on; found key => hanging
        dirs = {1: [(0, 1, 0), (0, 0, 1), (0, 1, 1)], 2: [(1, 0, 0), (0, 0, 1), (1, 0, 1)],
                3: [(0, 0, 1)], 4: [(1, 0, 0), (0, 1, 0), (1, 1, 0)], 5: [(0, 1, 0)], 6: [(1, 0, 0)]}
        self.node = []
        for o in self.octs:
            x, y, z, l = o
            hang = False
            if l > 1:
                cn = child_number(o, Dd)
                s = side(Dd, l)
                for (dx, dy, dz) in dirs.get(cn, []):
                    if (dx and not x) or (dy and not y) or (dz and not z):
                        continue
                    t = (x - dx * s, y - dy * s, z - dz * s, l)
                    if parent(t, Dd) in allset:
                        hang = True
            self.node.append(not hang)
        self.node_id = {}
        for o, isn in zip(self.octs, self.node):
            if isn:
                self.node_id[o[:3]] = len(self.node_id)
        self.N = len(self.node_id)
        # element node lists, hanging masks, child numbers, levels
        self.nlist = np.zeros((self.E, 8), np.uint32)
        self.hnmask = np.zeros(self.E, np.uint8)
        self.child = np.zeros(self.E, np.uint8)
        self.lev = np.zeros(self.E, np.uint8)
        for e, o in enumerate(elems):
            x, y, z, l = o
            s = side(Dd, l)
            p = parent(o, Dd)
            m = 0
            for k in range(8):
                q = (x + (k & 1) * s, y + ((k >> 1) & 1) * s, z + ((k >> 2) & 1) * s)
                if q in self.node_id:
                    self.nlist[e, k] = self.node_id[q]
                else:
                    m |= 1 << k
                    qp = (p[0] + (k & 1) * 2 * s, p[1] + ((k >> 1) & 1) * 2 * s, p[2] + ((k >> 2) & 1) * 2 * s)
                    self.nlist[e, k] = self.node_id[qp]
            self.hnmask[e] = m
            self.child[e] = child_number(o, Dd)
            self.lev[e] = l

    def node_positions(self):
        pos = np.zeros((self.N, 3), np.int64)
        for q, i in self.node_id.items():
            pos[i] = q
        return pos


def regular_nodes_geometric(da):
    """Independent check of the regular-node flags: an anchor is regular iff it is a vertex of
    every element whose closed box contains it."""
    out = []
    D = da.Dda
    for (x, y, z, l) in da.octs:
        reg = True
        for (ex, ey, ez, el) in da.elems:
            s = side(D, el)
            if ex <= x <= ex + s and ey <= y <= ey + s and ez <= z <= ez + s:
                if not ((x in (ex, ex + s)) and (y in (ey, ey + s)) and (z in (ez, ez + s))):
                    reg = False
                    break
        out.append(reg)
    return out
