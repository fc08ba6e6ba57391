"""H(curl) fields: Zaglmayr's hierarchical Nedelec-type functions
(src/shapefunction/hierarchical/hcurl/hcurl{line,triangle,quadrangle,
tetrahedron,hexahedron}.cpp), their covariant mapping (expression.cpp:30-35)
and the curl of an hcurl field (sl.cpp:1184-1222), pinned by what the
mathematics fixes:

- counts per element and order are the reference's count() formulas;
- the curl table is the curl of the value table (central differences), in
  every edge / face orientation;
- at order p >= 1 the functions are linearly independent and reproduce every
  vector polynomial of degree p (simplices: the full P_p^d; quadrangles and
  hexahedra contain it), in every orientation; order 0 is the Whitney space;
- the lowest-order functions have tangential moment +-1 on their own edge and
  0 on every other edge;
- L2 projection of a vector polynomial f of degree <= order over a mesh with
  randomly renumbered nodes (mixed edge and face orientations): x . b equals
  the exact integral of |f|^2 only if the projection reproduces f, which
  needs tangentially continuous functions across every shared edge and face;
  with the same DOF numbering, x^T K x of the curl-curl matrix equals the
  integral of |curl f|^2 (pins the curl tables and the J^T / det J mapping);
- -K + k^2 M assembled as one integral equals the two assembled apart
  (value and curl slices of one contribution).
GPU: the HIP assembly of the same problems against the oracle."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import sparselizard_amd as sa
from sparselizard_amd import _lib as L
from oracle import py_oracle

LINE, TRI, QUAD, TET, HEX = 1, 2, 3, 4, 5
DIM = {TRI: 2, QUAD: 2, TET: 3, HEX: 3}
ELEM = {"tri": TRI, "quad": QUAD, "tet": TET, "hex": HEX}
# number of total orientations to sample per type (edges, then faces in base 8)
ORIENTS = {TRI: [0, 1, 5, 6 + 8 * 3], TET: [0, 1, 37, 63 + 64 * (5 + 8 * 3 + 64 * 2 + 512 * 4)],
           QUAD: [0, 3, 9, 15 + 16 * 7], HEX: [0, 1, 4095, 4095 + 4096 * (7 + 8 * 3 + 64 * 5 + 512 * 1 + 4096 * 6 + 32768 * 2)]}


def refpts(t, n, rng):
    if t == TRI:
        u = rng.uniform(0, 1, (n, 2))
        u = np.where(u.sum(1, keepdims=True) > 1, 1 - u, u)
        return np.column_stack([u, np.zeros(n)])
    if t == TET:
        return rng.dirichlet([1.0, 1.0, 1.0, 1.0], n)[:, :3]
    if t == QUAD:
        return np.column_stack([rng.uniform(-1, 1, (n, 2)), np.zeros(n)])
    return rng.uniform(-1, 1, (n, 3))


def test_hcurl_counts():
    for p in range(0, 7):
        tri = 3 if p == 0 else 3 * (p + 1) + (p - 1) * (p + 1)
        tet = 6 if p == 0 else 6 * (p + 1) + 4 * (p - 1) * (p + 1) + (p - 2) * (p - 1) * (p + 1) // 2
        quad = 2 * p * (p + 1) + 4 * (p + 1)
        hexa = 3 * p * p * (p + 1) + 12 * p * (p + 1) + 12 * (p + 1)
        assert [L.hcurl_count(t, p) for t in (LINE, TRI, QUAD, TET, HEX)] == [p + 1, tri, quad, tet, hexa], p


@pytest.mark.parametrize("t", [TRI, QUAD, TET, HEX])
def test_hcurl_curl_table_is_curl_of_values(t):
    rng = np.random.default_rng(t)
    h = 1e-6
    for p in range(0, 4):
        pts = refpts(t, 6, rng)
        for o in ORIENTS[t]:
            c = L.hcurl_tables(t, p, pts, L.HCURL_CURL, o)[1:]
            g = np.zeros((3, 3) + c.shape[1:])
            for d in range(DIM[t]):
                e = np.zeros(3)
                e[d] = h
                g[:, d] = (L.hcurl_tables(t, p, pts + e, L.HCURL_VALUE, o)[1:] - L.hcurl_tables(t, p, pts - e, L.HCURL_VALUE, o)[1:]) / (2 * h)
            fd = np.stack([g[2, 1] - g[1, 2], g[0, 2] - g[2, 0], g[1, 0] - g[0, 1]])
            assert abs(c - fd).max() <= 1e-6 * max(1.0, abs(c).max()), (t, p, o)
            if DIM[t] == 2:
                assert not c[:2].any()


@pytest.mark.parametrize("t", [TRI, QUAD, TET, HEX])
def test_hcurl_spans_vector_polynomials(t):
    rng = np.random.default_rng(10 + t)
    dim = DIM[t]
    for p in range(1, 5 if dim == 2 else 4):
        nff = L.hcurl_count(t, p)
        n = max(120, 2 * nff)
        pts = refpts(t, n, rng)
        exps = [e for e in itertools.product(range(p + 1), repeat=dim) if sum(e) <= p]
        for o in ORIENTS[t][:3]:
            V = L.hcurl_tables(t, p, pts, L.HCURL_VALUE, o)[1:1 + dim]  # [comp][ff][pt]
            A = V.transpose(0, 2, 1).reshape(dim * n, nff)
            assert np.linalg.matrix_rank(A) == nff, (t, p, o)
            for c in range(dim):
                for e in exps:
                    tgt = np.zeros((dim, n))
                    tgt[c] = np.prod([pts[:, k] ** e[k] for k in range(dim)], axis=0)
                    coef = np.linalg.lstsq(A, tgt.reshape(-1), rcond=None)[0]
                    assert abs(A @ coef - tgt.reshape(-1)).max() <= 1e-9, (t, p, o, c, e)
        if dim == 2:
            assert not L.hcurl_tables(t, p, pts, L.HCURL_VALUE, 0)[3].any()


CORNERS = {TRI: [(0, 0, 0), (1, 0, 0), (0, 1, 0)], TET: [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
           QUAD: [(-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)],
           HEX: [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]}


@pytest.mark.parametrize("t", [TRI, QUAD, TET, HEX])
def test_hcurl_lowest_order_edge_moments(t):
    gx, gw = np.polynomial.legendre.leggauss(6)
    C = np.array(CORNERS[t], float)
    edges = [(a, b) for a, b in itertools.combinations(range(len(C)), 2)
             if t in (TRI, TET) or np.count_nonzero(C[a] - C[b]) == 1]
    for o in ORIENTS[t]:
        M = []
        for a, b in edges:
            pts = C[a] + np.outer((gx + 1) / 2, C[b] - C[a])
            V = L.hcurl_tables(t, 0, pts, L.HCURL_VALUE, o)[1:]
            M.append(np.einsum("c,cfp->fp", C[b] - C[a], V) @ gw / 2)
        M = np.abs(np.array(M))
        assert np.allclose(np.sort(M, axis=0)[-1], 1.0, atol=1e-13) and np.allclose(np.sort(M, axis=0)[:-1], 0.0, atol=1e-13)
        assert sorted(np.argmax(M, axis=0)) == list(range(len(edges)))


# ---- projection over meshes (oracle assembly of the formulation layer) ----

def fpoly(X, deg, is3d):
    """The problem's vector polynomial (hcurlprojection in problems.cpp) and its curl."""
    n = len(X)
    val = np.zeros((3, n))
    grad = np.zeros((3, 3, n))
    for comp in range(3 if is3d else 2):
        for a in range(deg + 1):
            for b in range(deg + 1 - a):
                for c in range(deg + 1 - a - b if is3d else 1):
                    co = (-1.0 if (comp + a + b + c) % 2 else 1.0) / (1.0 + a + 2 * b + 3 * c + 5 * comp)
                    x, y, z = X[:, 0], X[:, 1], X[:, 2]
                    val[comp] += co * x ** a * y ** b * z ** c
                    if a:
                        grad[comp, 0] += co * a * x ** (a - 1) * y ** b * z ** c
                    if b:
                        grad[comp, 1] += co * b * x ** a * y ** (b - 1) * z ** c
                    if c:
                        grad[comp, 2] += co * c * x ** a * y ** b * z ** (c - 1)
    curl = np.stack([grad[2, 1] - grad[1, 2], grad[0, 2] - grad[2, 0], grad[1, 0] - grad[0, 1]])
    return val, curl


def exact_integrals(prob, deg, is3d):
    d = prob.desc()
    num = prob.numeric()
    xyz = np.ctypeslib.as_array(num.node_xyz, shape=(3 * prob.n_nodes(),)).reshape(-1, 3)
    ff = cc = 0.0
    dim = 3 if is3d else 2
    for b in range(d.n_buckets):
        bk = d.buckets[b]
        pts, w = L.gauss_rule(bk.elem_type, 2 * deg + 4)
        G = L.geometry_tables(bk.elem_type, bk.n_geo, pts)
        geo = np.ctypeslib.as_array(bk.geo_nodes, shape=(bk.n_elem, bk.n_geo))
        xe = xyz[geo]                                     # [elem][geo][3]
        X = np.einsum("gp,egc->epc", G[0], xe)
        J = np.stack([np.einsum("gp,egc->epc", G[1 + r], xe) for r in range(dim)], 2)[..., :dim]
        det = np.abs(np.linalg.det(J))
        for e in range(bk.n_elem):
            v, c = fpoly(X[e], deg, is3d)
            ff += np.sum(w * det[e] * (v ** 2).sum(0))
            cc += np.sum(w * det[e] * (c ** 2).sum(0))
    return ff, cc


def oracle_system(kind, spec):
    p = sa.Problem(None, kind, spec + ";meshpipe=host")
    info = p.info()
    o = py_oracle.assemble(p.desc_ptr(), p.numeric(), info["ndofs"], info["nfree"])
    nf = info["nfree"]
    return p, sp.csr_matrix((o.Aval, o.Acol, o.Arow), shape=(nf, nf)), o.rhs


CASES = [("tri", o) for o in (0, 1, 2, 3)] + [("quad", o) for o in (0, 1, 2, 3)] + \
        [("tet", o) for o in (0, 1, 2)] + [("hex", o) for o in (0, 1, 2)]


@pytest.mark.parametrize("elem,order", CASES)
def test_hcurl_projection_reproduces_polynomials(elem, order):
    is3d = elem in ("tet", "hex")
    spec = "elem=%s;n=%d;order=%d;deg=%d" % (elem, 2 if is3d else 3, order, order)
    p, M, b = oracle_system("hcurlprojection", spec + ";term=mass")
    _, K, b2 = oracle_system("hcurlprojection", spec + ";term=curl")
    assert np.array_equal(b, b2)
    assert abs(M - M.T).max() <= 1e-14 * abs(M).max() and abs(K - K.T).max() <= 1e-13 * max(abs(K).max(), 1e-300)
    x = spl.spsolve(M.tocsc(), b)
    ff, cc = exact_integrals(p, order, is3d)
    assert abs(x @ b - ff) <= 1e-10 * ff, (x @ b, ff)
    assert abs(x @ K @ x - cc) <= 1e-10 * max(cc, 1e-3), (x @ K @ x, cc)
    # planted check of the test itself: one degree more is not reproduced
    # (one cell per direction, so that the defect is far above the tolerance)
    if order >= 1:
        p3, M3, b3 = oracle_system("hcurlprojection", "elem=%s;n=1;order=%d;deg=%d;term=mass" % (elem, order, order + 1))
        x3 = spl.spsolve(M3.tocsc(), b3)
        ff3, _ = exact_integrals(p3, order + 1, is3d)
        assert ff3 - x3 @ b3 > 1e-7 * ff3, (ff3, x3 @ b3)


@pytest.mark.parametrize("elem,order", [("tri", 2), ("quad", 1), ("tet", 1), ("hex", 1)])
def test_hcurl_maxwell_is_curl_and_mass_slices(elem, order):
    is3d = elem in ("tet", "hex")
    spec = "elem=%s;n=%d;order=%d" % (elem, 2 if is3d else 3, order)
    _, M, _ = oracle_system("hcurlprojection", spec + ";term=mass")
    _, K, _ = oracle_system("hcurlprojection", spec + ";term=curl")
    _, A, _ = oracle_system("hcurlprojection", spec + ";term=maxwell;k2=7")
    ref = -K + 7.0 * M
    assert abs(A - ref).max() <= 1e-12 * abs(ref).max()


@pytest.mark.parametrize("elem,order", [("quad", 3), ("hex", 2)])
def test_hcurl_orientation_folding(elem, order):
    # quadrangle and hexahedron functions fold to orientation 0 (signed
    # renumbering): one bucket instead of one per orientation, same matrix;
    # triangle faces (order >= 2) do not fold: tetrahedra keep their batches
    spec = "elem=%s;n=%d;order=%d;term=maxwell" % (elem, 3 if elem == "quad" else 2, order)
    pf, Af, bf = oracle_system("hcurlprojection", spec)
    pu, Au, bu = oracle_system("hcurlprojection", spec + ";fold=0")
    assert pf.info()["nbuckets"] == 1 and pu.info()["nbuckets"] > 4
    assert abs(Af - Au).max() <= 1e-13 * abs(Au).max()
    assert abs(bf - bu).max() <= 1e-13 * abs(bu).max()
    pt, _, _ = oracle_system("hcurlprojection", "elem=tet;n=2;order=2;term=mass")
    assert pt.info()["nbuckets"] > 1


def test_hcurl_rejects_unsupported_uses():
    with pytest.raises(sa.SlxError):
        sa.Problem(None, "hcurlprojection", "elem=tri;n=2;order=-1;meshpipe=host")


# ---- GPU: HIP assembly against the oracle ----

GPU_CASES = [("tri", 2, "maxwell"), ("tri", 0, "mass"), ("quad", 2, "maxwell"), ("tet", 2, "maxwell"),
             ("tet", 3, "curl"), ("hex", 1, "maxwell"), ("hex", 2, "mass"), ("hex", 3, "maxwell")]


@pytest.mark.gpu
@pytest.mark.parametrize("elem,order,term", GPU_CASES)
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("f16split", 1e-6)])
def test_hcurl_gpu_matches_oracle(ctx, elem, order, term, precision, tol):
    is3d = elem in ("tet", "hex")
    spec = "elem=%s;n=%d;order=%d;term=%s" % (elem, 2 if is3d else 4, order, term)
    print("case hcurl", spec, precision, flush=True)
    prob = sa.Problem(ctx, "hcurlprojection", spec)
    info = prob.info()
    got = prob.generate(precision).csr(info["ndofs"])
    ref = py_oracle.assemble(prob.desc_ptr(), prob.numeric(), info["ndofs"], info["nfree"])
    for name in ("Arow", "Acol", "Drow", "Dcol"):
        assert np.array_equal(getattr(got, name), getattr(ref, name)), name
    for name in ("Aval", "rhs"):
        a, b = getattr(got, name), getattr(ref, name)
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert err <= tol, (name, err)


@pytest.mark.gpu
@pytest.mark.parametrize("elem,order", [("tri", 3), ("hex", 2)])
def test_hcurl_gpu_projection_reproduces_polynomials(ctx, elem, order):
    # the GPU matrix itself reproduces f (FP64 assembly, host solve)
    is3d = elem in ("tet", "hex")
    spec = "elem=%s;n=%d;order=%d;deg=%d;term=mass" % (elem, 2 if is3d else 3, order, order)
    prob = sa.Problem(ctx, "hcurlprojection", spec)
    info = prob.info()
    got = prob.generate("fp64").csr(info["ndofs"])
    nf = info["nfree"]
    M = sp.csr_matrix((got.Aval, got.Acol, got.Arow), shape=(nf, nf))
    x = spl.spsolve(M.tocsc(), got.rhs)
    ff, _ = exact_integrals(prob, order, is3d)
    assert abs(x @ got.rhs - ff) <= 1e-10 * ff
