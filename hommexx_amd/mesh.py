"""Equiangular gnomonic cubed-sphere mesh with np=4 GLL elements.

This module plays the role of the Fortran host's mesh setup (cube_mod.F90,
gridgraph_mod.F90, mass_matrix_mod.F90 in the reference): it produces the
per-element metric arrays handed to ``init_elements_2d_c`` and the element
connections handed to ``add_connection``.  It is host-side set-up code, not
part of the time step.

Conventions (shared with the C-ABI and the kernels):

* Element-local GLL indices follow the Fortran arrays ``(i, j)`` with ``i``
  along the first reference coordinate ``a`` and ``j`` along ``b``.  The flat
  point index is ``pt = 4*(j-1) + (i-1)`` (Fortran column-major), which is the
  reference C++ ``(igp, jgp) = (j-1, i-1)`` with ``jgp`` fastest.
* ``D`` maps reference-element contravariant components to the (zonal,
  meridional) velocity on the unit sphere: ``D[r, c] = e_r . dX/d(ref_c)``
  with ``e_0`` eastward, ``e_1`` northward, ``ref_0 = a``, ``ref_1 = b``
  (cube_mod.F90:215-454 ``metric_atomic`` / ``dmap_equiangular``).
* Connection positions use the Fortran numbering passed through
  ``add_connection``: W=1, E=2, S=3, N=4, SW=5, SE=6, NW=7, NE=8
  (converted to S,N,W,E,SW,SE,NW,NE by mpi_cxx_f90_interface.cpp:44-49).
  "South" is the ``b = -1`` edge, "west" the ``a = -1`` edge.
* Every face frame ``(e1, e2, n)`` is right-handed, so two elements that share
  an edge traverse it in opposite counter-clockwise senses; this is what makes
  the reference's position-pair direction table
  (ConnectivityHelpers.hpp:133-142) valid.
"""
import math

import numpy as np

NP = 4
REARTH = 6.376e6
OMEGA = 7.292e-5

# Fortran connection positions (1-based) and their reference-C++ counterparts.
F_W, F_E, F_S, F_N, F_SW, F_SE, F_NW, F_NE = 1, 2, 3, 4, 5, 6, 7, 8

# face frames: (normal, e1, e2), right-handed: e1 x e2 = n
_FACES = [
    ((1, 0, 0), (0, 1, 0), (0, 0, 1)),
    ((0, 1, 0), (-1, 0, 0), (0, 0, 1)),
    ((-1, 0, 0), (0, -1, 0), (0, 0, 1)),
    ((0, -1, 0), (1, 0, 0), (0, 0, 1)),
    ((0, 0, 1), (0, 1, 0), (-1, 0, 0)),
    ((0, 0, -1), (0, 1, 0), (1, 0, 0)),
]


def gll_points_weights():
    """np=4 Gauss-Lobatto-Legendre nodes on [-1, 1] and their weights."""
    s = 1.0 / math.sqrt(5.0)
    x = np.array([-1.0, -s, s, 1.0])
    w = np.array([1.0 / 6.0, 5.0 / 6.0, 5.0 / 6.0, 1.0 / 6.0])
    return x, w


def derivative_matrix():
    """``dvv[a, b] = L_b'(x_a)``: derivative of GLL basis ``b`` at node ``a``.

    This is the C++ layout (Derivative.cpp:20-32 copies the Fortran ``Dvv``
    flat, so C ``dvv(a, b)`` equals Fortran ``Dvv(b+1, a+1)``).  The Fortran
    array itself is ``dvv.T`` in column-major order, i.e. ``dvv`` flattened in
    C order is exactly what ``init_derivative_c`` receives.
    """
    x, _ = gll_points_weights()
    n = NP - 1
    # Legendre P_3 at the nodes
    p3 = 0.5 * (5.0 * x ** 3 - 3.0 * x)
    d = np.zeros((NP, NP))
    for a in range(NP):
        for b in range(NP):
            if a != b:
                d[a, b] = p3[a] / (p3[b] * (x[a] - x[b]))
    d[0, 0] = -n * (n + 1) / 4.0
    d[NP - 1, NP - 1] = n * (n + 1) / 4.0
    return d


def _hilbert_index(n, x, y):
    """Position of cell (x, y) along the Hilbert curve over an n x n grid
    (n a power of two)."""
    d = 0
    s = n // 2
    while s > 0:
        rx = 1 if x & s else 0
        ry = 1 if y & s else 0
        d += s * s * ((3 * rx) ^ ry)
        if ry == 0:
            if rx == 1:
                x, y = s - 1 - x, s - 1 - y
            x, y = y, x
        s //= 2
    return d


def sfc_order(ne):
    """Face-major element order with each face's ne x ne elements along a
    Hilbert curve (cells of the enclosing power-of-two grid outside the face
    are skipped), like HOMME's space-filling-curve element ordering
    (src/share/spacecurve_mod.F90): elements that share an edge are mostly
    close in the order.  Returns the natural (face, ej, ei) row-major index
    of each position."""
    n = 1
    while n < ne:
        n *= 2
    face = sorted((_hilbert_index(n, i, j), j * ne + i) for j in range(ne) for i in range(ne))
    return np.array([f * ne * ne + c for f in range(6) for _, c in face], dtype=np.int64)


class CubedSphere:
    """Geometry and connectivity of an ``ne x ne x 6`` cubed sphere.

    ``order``: "sfc" (default) numbers the elements along a space-filling
    curve on each face, as HOMME does; "rowmajor" numbers them face by face,
    row by row."""

    def __init__(self, ne, order="sfc"):
        if ne < 2:
            raise ValueError("ne must be >= 2 (each element needs at most one missing corner)")
        if order not in ("sfc", "rowmajor"):
            raise ValueError("order must be 'sfc' or 'rowmajor'")
        self.ne = ne
        self.nelem = 6 * ne * ne
        self.order = sfc_order(ne) if order == "sfc" else np.arange(self.nelem)
        self._rank = np.empty_like(self.order)
        self._rank[self.order] = np.arange(self.nelem)
        self._build_geometry()
        self._build_connectivity()
        self._build_mass()

    # ------------------------------------------------------------------ geometry
    def elem_index(self, face, ei, ej):
        return int(self._rank[face * self.ne * self.ne + ej * self.ne + ei])

    def _build_geometry(self):
        ne = self.ne
        x, w = gll_points_weights()
        dang = 0.5 * math.pi / ne
        nelem = self.nelem
        # per element, per point (j, i): 3-D position and its reference derivatives
        ei = np.arange(ne)
        pos = np.zeros((nelem, NP, NP, 3))
        dpda = np.zeros((nelem, NP, NP, 3))
        dpdb = np.zeros((nelem, NP, NP, 3))
        corner3d = np.zeros((nelem, 4, 3))  # SW, SE, NW, NE
        self.face_of = np.repeat(np.arange(6), ne * ne)
        # element-local angles: alpha[ei, i], beta[ej, j]
        al = -0.25 * math.pi + (ei[:, None] + 0.5 * (1.0 + x[None, :])) * dang
        be = al.copy()
        ca = np.array([-0.25 * math.pi + ei * dang, -0.25 * math.pi + (ei + 1) * dang]).T  # corners (ei, 2)
        for f, (n, e1, e2) in enumerate(_FACES):
            n = np.array(n, float)
            e1 = np.array(e1, float)
            e2 = np.array(e2, float)
            # grids over (ej, j, ei, i)
            A = al[None, None, :, :]
            B = be[:, :, None, None]
            ta, tb = np.tan(A), np.tan(B)
            X = n + ta[..., None] * e1 + tb[..., None] * e2
            rn = np.sqrt(np.sum(X * X, axis=-1))[..., None]
            r = X / rn
            dXa = (e1 / np.cos(A)[..., None] ** 2) * np.ones_like(tb)[..., None]
            dXb = (e2 / np.cos(B)[..., None] ** 2) * np.ones_like(ta)[..., None]
            dra = (dXa - r * np.sum(r * dXa, axis=-1)[..., None]) / rn
            drb = (dXb - r * np.sum(r * dXb, axis=-1)[..., None]) / rn
            sl = slice(f * ne * ne, (f + 1) * ne * ne)
            # (ej, j, ei, i) -> element index ej*ne + ei, point (j, i)
            to_elem = lambda a: np.transpose(a, (0, 2, 1, 3, 4)).reshape(ne * ne, NP, NP, 3)
            pos[sl] = to_elem(r)
            dpda[sl] = to_elem(0.5 * dang * dra)
            dpdb[sl] = to_elem(0.5 * dang * drb)
            for k, (ii, jj) in enumerate(((0, 0), (1, 0), (0, 1), (1, 1))):
                AA = ca[:, ii][None, :]
                BB = ca[:, jj][:, None]
                Xc = n + np.tan(AA)[..., None] * e1 + np.tan(BB)[..., None] * e2
                Xc = Xc / np.sqrt(np.sum(Xc * Xc, axis=-1))[..., None]
                corner3d[sl, k] = Xc.reshape(ne * ne, 3)
        # element numbering: position i holds the natural element order[i]
        pos, dpda, dpdb, corner3d = pos[self.order], dpda[self.order], dpdb[self.order], corner3d[self.order]
        self.face_of = self.face_of[self.order]
        # snap near-zero coordinates (the poles and face centres are nodes when ne is even)
        pos[np.abs(pos) < 1e-14] = 0.0
        self.pos = pos
        self.corner3d = corner3d
        lat = np.arcsin(np.clip(pos[..., 2], -1.0, 1.0))
        lon = np.arctan2(pos[..., 1], pos[..., 0])
        pole = (np.abs(pos[..., 0]) < 1e-14) & (np.abs(pos[..., 1]) < 1e-14)
        lon = np.where(pole, 0.0, lon)
        lon = np.mod(lon, 2.0 * math.pi)
        self.lat = lat
        self.lon = lon
        sl, cl = np.sin(lon), np.cos(lon)
        sp, cp = np.sin(lat), np.cos(lat)
        elon = np.stack([-sl, cl, np.zeros_like(sl)], axis=-1)
        elat = np.stack([-sp * cl, -sp * sl, cp], axis=-1)
        D = np.zeros((nelem, NP, NP, 2, 2))
        D[..., 0, 0] = np.sum(elon * dpda, axis=-1)
        D[..., 1, 0] = np.sum(elat * dpda, axis=-1)
        D[..., 0, 1] = np.sum(elon * dpdb, axis=-1)
        D[..., 1, 1] = np.sum(elat * dpdb, axis=-1)
        det = D[..., 0, 0] * D[..., 1, 1] - D[..., 0, 1] * D[..., 1, 0]
        Dinv = np.zeros_like(D)
        Dinv[..., 0, 0] = D[..., 1, 1] / det
        Dinv[..., 0, 1] = -D[..., 0, 1] / det
        Dinv[..., 1, 0] = -D[..., 1, 0] / det
        Dinv[..., 1, 1] = D[..., 0, 0] / det
        met = np.einsum("...ki,...kj->...ij", D, D)
        metinv = np.zeros_like(met)
        metinv[..., 0, 0] = met[..., 1, 1] / (det * det)
        metinv[..., 0, 1] = -met[..., 0, 1] / (det * det)
        metinv[..., 1, 0] = -met[..., 1, 0] / (det * det)
        metinv[..., 1, 1] = met[..., 0, 0] / (det * det)
        self.D = D  # (nelem, j, i, r, c)
        self.Dinv = Dinv
        self.metdet = np.abs(det)
        self.metinv = metinv
        self.mp = np.broadcast_to(np.outer(w, w), (nelem, NP, NP)).copy()  # (j, i)
        self.fcor = 2.0 * OMEGA * np.sin(lat)
        # vec_sphere2cart(i,j,xyz,comp): comp 0 = zonal, 1 = meridional
        self.vec_sph2cart = np.stack([elon, elat], axis=-1)  # (nelem, j, i, xyz, comp)

    # -------------------------------------------------------------- connectivity
    def _build_connectivity(self):
        nelem = self.nelem
        key = lambda p: tuple(np.round(p * 1e9).astype(np.int64))
        node_id = {}
        cid = np.zeros((nelem, 4), dtype=np.int64)  # SW, SE, NW, NE
        for ie in range(nelem):
            for k in range(4):
                kk = key(self.corner3d[ie, k])
                if kk not in node_id:
                    node_id[kk] = len(node_id)
                cid[ie, k] = node_id[kk]
        self.corner_node = cid
        # edges as (corner pair) in the element's natural order
        edge_corners = {F_S: (0, 1), F_N: (2, 3), F_W: (0, 2), F_E: (1, 3)}
        corner_pos = {0: F_SW, 1: F_SE, 2: F_NW, 3: F_NE}
        edge_map = {}
        for ie in range(nelem):
            for pos, (c0, c1) in edge_corners.items():
                k = frozenset((cid[ie, c0], cid[ie, c1]))
                edge_map.setdefault(k, []).append((ie, pos))
        node_elems = {}
        for ie in range(nelem):
            for k in range(4):
                node_elems.setdefault(cid[ie, k], []).append((ie, k))
        conns = []  # (lid, gid, pos, pid, lid2, gid2, pos2, pid2), 1-based Fortran convention
        # edge neighbours
        for k, lst in edge_map.items():
            assert len(lst) == 2, "every cube edge must be shared by exactly two elements"
            (e1, p1), (e2, p2) = lst
            conns.append((e1, p1, e2, p2))
            conns.append((e2, p2, e1, p1))
        # corner neighbours: elements sharing exactly one node and no edge
        for ie in range(nelem):
            edge_nbrs = set()
            for pos, (c0, c1) in edge_corners.items():
                k = frozenset((cid[ie, c0], cid[ie, c1]))
                for (e, p) in edge_map[k]:
                    if e != ie:
                        edge_nbrs.add(e)
            for k in range(4):
                cands = [(e, kk) for (e, kk) in node_elems[cid[ie, k]] if e != ie and e not in edge_nbrs]
                assert len(cands) <= 1
                if cands:
                    e, kk = cands[0]
                    conns.append((ie, corner_pos[k], e, corner_pos[kk]))
        conns.sort()
        self.connections = [(a + 1, a + 1, p, 1, b + 1, b + 1, q, 1) for (a, p, b, q) in conns]
        # neighbour table in Fortran positions, for host-side checks: nbr[ie, fpos-1] = (elem, fpos2)
        nbr = -np.ones((nelem, 8, 2), dtype=np.int64)
        for (a, p, b, q) in conns:
            nbr[a, p - 1] = (b, q)
        self.nbr = nbr
        self.num_unique_nodes = None

    # ------------------------------------------------------------------- masses
    def dss_scalar_reference(self, f):
        """Assemble a 2-D field ``f[ie, j, i]`` by summing all copies of every
        GLL node (geometric matching); used only to build ``rspheremp``."""
        key = np.round(self.pos * 1e9).astype(np.int64)
        flat = key.reshape(-1, 3)
        _, inv = np.unique(flat, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        sums = np.zeros(inv.max() + 1)
        np.add.at(sums, inv, f.reshape(-1))
        self.num_unique_nodes = len(sums)
        return sums[inv].reshape(f.shape)

    def _build_mass(self):
        self.spheremp = self.mp * self.metdet
        self.rspheremp = 1.0 / self.dss_scalar_reference(self.spheremp)

    # -------------------------------------------------------- Fortran-layout out
    def fortran_arrays(self):
        """2-D metric arrays in the flat Fortran layouts ``init_elements_2d_c``
        receives (cxx_f90_interface.cpp:224-233, Elements.cpp:81-190).

        Returned as C-contiguous numpy arrays whose memory order equals the
        Fortran column-major order, e.g. ``D`` has numpy shape
        ``(nelem, 2[c], 2[r], 4[j], 4[i])`` for Fortran ``D(np,np,2,2,nelemd)``.
        """
        tr = lambda m: np.ascontiguousarray(np.transpose(m, (0, 4, 3, 1, 2)))  # (e,j,i,r,c)->(e,c,r,j,i)
        out = {
            "D": tr(self.D),
            "Dinv": tr(self.Dinv),
            "metinv": tr(self.metinv),
            "fcor": np.ascontiguousarray(self.fcor),
            "mp": np.ascontiguousarray(self.mp),
            "spheremp": np.ascontiguousarray(self.spheremp),
            "rspheremp": np.ascontiguousarray(self.rspheremp),
            "metdet": np.ascontiguousarray(self.metdet),
            # vec_sphere2cart(np,np,3,2): (e, comp, xyz, j, i)
            "vec_sph2cart": np.ascontiguousarray(np.transpose(self.vec_sph2cart, (0, 4, 3, 1, 2))),
            # constant-coefficient HV only: tensor viscosity is the identity
            "tensorvisc": np.ascontiguousarray(
                np.broadcast_to(np.eye(2)[None, :, :, None, None], (self.nelem, 2, 2, NP, NP))),
        }
        return out
