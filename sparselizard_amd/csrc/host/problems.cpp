// Synthetic benchmark workloads (see problems.h).
#include "problems.h"
#include <numeric>
#include <cmath>
#include <sstream>

namespace slx {

static uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

Rng::Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& s : s_) s = splitmix64(x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t Rng::next() {
    uint64_t r = rotl(s_[1] * 5, 7) * 9;
    uint64_t t = s_[1] << 17;
    s_[2] ^= s_[0];
    s_[3] ^= s_[1];
    s_[1] ^= s_[2];
    s_[0] ^= s_[3];
    s_[2] ^= t;
    s_[3] = rotl(s_[3], 45);
    return r;
}

double Rng::uniform() { return (double)(next() >> 11) * 0x1.0p-53; }

int Rng::uniform_int(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }

double Rng::normal() {
    if (has_spare_) {
        has_spare_ = false;
        return spare_;
    }
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    double r = std::sqrt(-2.0 * std::log(u1)), th = 2.0 * M_PI * u2;
    spare_ = r * std::sin(th);
    has_spare_ = true;
    return r * std::cos(th);
}

std::map<std::string, std::string> parse_spec(const std::string& spec) {
    std::map<std::string, std::string> out;
    std::stringstream ss(spec);
    std::string item;
    while (std::getline(ss, item, ';')) {
        auto eq = item.find('=');
        if (eq == std::string::npos) continue;
        out[item.substr(0, eq)] = item.substr(eq + 1);
    }
    return out;
}

namespace {

double getd(const std::map<std::string, std::string>& s, const char* k, double dflt) {
    auto it = s.find(k);
    return it == s.end() ? dflt : std::stod(it->second);
}
int geti(const std::map<std::string, std::string>& s, const char* k, int dflt) {
    auto it = s.find(k);
    return it == s.end() ? dflt : std::stoi(it->second);
}

// Unit-square grid of n x n cells, lexicographic vertices, interior jitter.
std::vector<double> square_nodes(int n, double jitter, Rng& rng, double x0 = 0, double y0 = 0) {
    std::vector<double> xyz;
    double h = 1.0 / n;
    xyz.reserve(3 * (size_t)(n + 1) * (n + 1));
    for (int j = 0; j <= n; j++)
        for (int i = 0; i <= n; i++) {
            double x = i * h, y = j * h;
            if (i > 0 && i < n && j > 0 && j < n && jitter > 0) {
                x += rng.uniform(-jitter * h, jitter * h);
                y += rng.uniform(-jitter * h, jitter * h);
            }
            xyz.insert(xyz.end(), {x0 + x, y0 + y, 0.0});
        }
    return xyz;
}

void square_boundary(Mesh& m, int n, int base, int physreg) {
    auto v = [&](int i, int j) { return base + i + (n + 1) * j; };
    for (int i = 0; i < n; i++) m.add_element(LINE, {v(i, 0), v(i + 1, 0)}, physreg);
    for (int j = 0; j < n; j++) m.add_element(LINE, {v(n, j), v(n, j + 1)}, physreg);
    for (int i = n; i > 0; i--) m.add_element(LINE, {v(i, n), v(i - 1, n)}, physreg);
    for (int j = n; j > 0; j--) m.add_element(LINE, {v(0, j), v(0, j - 1)}, physreg);
}

void square_tris(Mesh& m, int n, int base, int physreg) {
    auto v = [&](int i, int j) { return base + i + (n + 1) * j; };
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            m.add_element(TRI, {v(i, j), v(i + 1, j), v(i + 1, j + 1)}, physreg);
            m.add_element(TRI, {v(i, j), v(i + 1, j + 1), v(i, j + 1)}, physreg);
        }
}

// Config: 2D Poisson, P1 triangles, u = 0 on the boundary, f constant.
std::shared_ptr<Problem> poisson2d(const std::map<std::string, std::string>& s) {
    auto p = std::make_shared<Problem>();
    int n = geti(s, "n", 200), order = geti(s, "order", 1);
    double jit = getd(s, "jitter", 0.2), f = getd(s, "f", 1.0);
    Rng rng((uint64_t)geti(s, "seed", 0));
    p->mesh = std::make_shared<Mesh>();
    Mesh& m = *p->mesh;
    m.set_nodes(square_nodes(n, jit, rng));
    square_boundary(m, n, 0, 2);
    square_tris(m, n, 0, 1);
    m.finalize();
    field u(m, "h1", "u");
    if (order != 1) u.setorder(1, order);
    u.setconstraint(2);
    p->fields = {u};
    p->form = std::make_shared<formulation>(m);
    using namespace sl;
    *p->form += integral(1, grad(dof(u)) * grad(tf(u)));
    *p->form += integral(1, -expression(f) * tf(u));
    std::ostringstream d;
    d << "2D Poisson P" << order << ", " << n << "x" << n << "x2 triangles, jitter " << jit;
    p->describe = d.str();
    return p;
}

// Steady convection-diffusion on the unit square (a non-symmetric system):
//   integral(grad(dof(u)) . grad(tf(u)) + (b . grad(dof(u))) tf(u) - f tf(u)),
// b = (bx, by) constant, u = 0 on the boundary; the geometry and numbering of
// poisson2d.
std::shared_ptr<Problem> convdiff2d(const std::map<std::string, std::string>& s) {
    auto p = std::make_shared<Problem>();
    int n = geti(s, "n", 64), order = geti(s, "order", 1);
    double jit = getd(s, "jitter", 0.2), f = getd(s, "f", 1.0), bx = getd(s, "bx", 20.0), by = getd(s, "by", 10.0);
    Rng rng((uint64_t)geti(s, "seed", 0));
    p->mesh = std::make_shared<Mesh>();
    Mesh& m = *p->mesh;
    m.set_nodes(square_nodes(n, jit, rng));
    square_boundary(m, n, 0, 2);
    square_tris(m, n, 0, 1);
    m.finalize();
    field u(m, "h1", "u");
    if (order != 1) u.setorder(1, order);
    u.setconstraint(2);
    p->fields = {u};
    p->form = std::make_shared<formulation>(m);
    using namespace sl;
    const expression b(1, 2, {expression(bx), expression(by)});
    *p->form += integral(1, grad(dof(u)) * grad(tf(u)) + (b * grad(dof(u))) * tf(u));
    *p->form += integral(1, -expression(f) * tf(u));
    std::ostringstream d;
    d << "2D convection-diffusion P" << order << ", b = (" << bx << ", " << by << "), " << n << "x" << n << "x2 triangles";
    p->describe = d.str();
    return p;
}

// Config: nonlinear 2D magnetostatics-type problem with Newton re-assembly.
// P2 field a on n x n x 2 jittered triangles of the unit square (a = 0 on the
// boundary), reluctivity nu(|grad a|) = 1 + 500 / (1 + |grad a|^2), source f.
// The weak form is linearized at the current field a0 as in the reference's
// saturation example (examples/magnetostatics-vector-potential-saturation-2d/
// main.cpp:93-94):  integral((h(b0) + dh/db(b0) (grad(dof a) - b0)) . grad(tf a))
// with b0 = grad a0, h = nu b0, dh/db = nu I + (nu'/|b0|) b0 b0^T and
// nu'/|b| = -1000 / (1 + |b|^2)^2; plus integral(-f tf a).  Each generate()
// re-assembles the Jacobian and the rhs at the field data last set.
std::shared_ptr<Problem> newton2d(const std::map<std::string, std::string>& s) {
    auto p = std::make_shared<Problem>();
    int n = geti(s, "n", 708), order = geti(s, "order", 2);
    double jit = getd(s, "jitter", 0.2), f = getd(s, "f", 500.0);
    Rng rng((uint64_t)geti(s, "seed", 4));
    p->mesh = std::make_shared<Mesh>();
    Mesh& m = *p->mesh;
    m.set_nodes(square_nodes(n, jit, rng));
    square_boundary(m, n, 0, 2);
    square_tris(m, n, 0, 1);
    m.finalize();
    field a(m, "h1", "a");
    a.setorder(1, order);
    a.setconstraint(2);
    p->fields = {a};
    p->form = std::make_shared<formulation>(m);
    using namespace sl;
    expression g = grad(a);
    expression n2 = g * g;
    expression one(1.0);
    expression nu = one + expression(500.0) / (one + n2);
    expression c = expression(-1000.0) / ((one + n2) * (one + n2));
    expression gd = grad(dof(a)), gt = grad(tf(a));
    expression dg = gd - g;
    *p->form += integral(1, (nu * g + nu * dg + c * (g * dg) * g) * gt);
    *p->form += integral(1, -expression(f) * tf(a));
    std::ostringstream d;
    d << "2D Newton magnetostatics P" << order << ", " << n << "x" << n << "x2 triangles, jitter " << jit << ", f " << f;
    p->describe = d.str();
    return p;
}

// Config: 3D isotropic elasticity, h1xyz order p on Kuhn tets of an n^3 grid,
// E per element lognormal(0, sigma), nu, clamped on x = 0, body force -z.
std::shared_ptr<Problem> elasticity3d(const std::map<std::string, std::string>& s) {
    auto p = std::make_shared<Problem>();
    int n = geti(s, "n", 64), order = geti(s, "order", 2);
    double jit = getd(s, "jitter", 0.15), sigma = getd(s, "sigma", 0.5), nu = getd(s, "nu", 0.3);
    Rng rng((uint64_t)geti(s, "seed", 1));
    p->mesh = std::make_shared<Mesh>();
    Mesh& m = *p->mesh;
    double h = 1.0 / n;
    int n1 = n + 1;
    std::vector<double> xyz;
    xyz.reserve(3 * (size_t)n1 * n1 * n1);
    for (int k = 0; k <= n; k++)
        for (int j = 0; j <= n; j++)
            for (int i = 0; i <= n; i++) {
                double c[3] = {i * h, j * h, k * h};
                bool interior = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
                if (interior && jit > 0)
                    for (int d = 0; d < 3; d++) c[d] += rng.uniform(-jit * h, jit * h);
                xyz.insert(xyz.end(), {c[0], c[1], c[2]});
            }
    m.set_nodes(xyz);
    auto v = [&](int i, int j, int k) { return i + n1 * (j + n1 * k); };
    // clamped face x = 0, split along the Kuhn face diagonal (j,k)-(j+1,k+1)
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++) {
            m.add_element(TRI, {v(0, j, k), v(0, j + 1, k), v(0, j + 1, k + 1)}, 2);
            m.add_element(TRI, {v(0, j, k), v(0, j, k + 1), v(0, j + 1, k + 1)}, 2);
        }
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++)
                for (auto& pm : perms) {
                    int c[3] = {i, j, k};
                    std::vector<int> tet = {v(c[0], c[1], c[2])};
                    for (int a = 0; a < 2; a++) {
                        c[pm[a]]++;
                        tet.push_back(v(c[0], c[1], c[2]));
                    }
                    tet.push_back(v(i + 1, j + 1, k + 1));
                    m.add_element(TET, tet, 1);
                }
    m.finalize();
    int ntet = m.count(TET);
    parameter E;
    std::vector<double> ev(ntet);
    for (int e = 0; e < ntet; e++) ev[e] = std::exp(sigma * rng.normal());
    E.setelementvalues(ev);
    field u(m, "h1xyz", "u");
    u.setorder(1, order);
    if (geti(s, "clamp", 1)) u.setconstraint(2);
    p->fields = {u};
    p->params = {E};
    p->form = std::make_shared<formulation>(m);
    using namespace sl;
    *p->form += integral(1, predefinedelasticity(dof(u), tf(u), E, expression(nu)));
    expression f(3, 1, {expression(0.0), expression(0.0), expression(-1.0)});
    *p->form += integral(1, f * tf(u));
    std::ostringstream d;
    d << "3D elasticity h1xyz P" << order << ", " << n << "^3 x 6 tets (" << ntet << "), jitter " << jit;
    p->describe = d.str();
    return p;
}

// Config: many independent quad instances, Q-order Lagrange-equivalent H1,
// K + k^2 M with k^2 uniform per instance, assembled as one block-diagonal set.
std::shared_ptr<Problem> quadbatch(const std::map<std::string, std::string>& s) {
    auto p = std::make_shared<Problem>();
    int ninst = geti(s, "instances", 512), nmin = geti(s, "nmin", 16), nmax = geti(s, "nmax", 160);
    int order = geti(s, "order", 3);
    Rng rng((uint64_t)geti(s, "seed", 3));
    std::vector<int> sizes(ninst);
    std::vector<double> k2(ninst);
    for (int i = 0; i < ninst; i++) {
        sizes[i] = rng.uniform_int(nmin, nmax);
        k2[i] = rng.uniform(1.0, 50.0);
    }
    p->mesh = std::make_shared<Mesh>();
    Mesh& m = *p->mesh;
    std::vector<double> xyz;
    for (int i = 0; i < ninst; i++) {
        std::vector<double> sq = square_nodes(sizes[i], 0.0, rng, 2.0 * i, 0.0);
        xyz.insert(xyz.end(), sq.begin(), sq.end());
    }
    m.set_nodes(xyz);
    std::vector<double> kelem;
    int base = 0;
    for (int i = 0; i < ninst; i++) {
        int n = sizes[i];
        auto v = [&](int a, int b) { return base + a + (n + 1) * b; };
        for (int j = 0; j < n; j++)
            for (int a = 0; a < n; a++) {
                m.add_element(QUAD, {v(a, j), v(a + 1, j), v(a + 1, j + 1), v(a, j + 1)}, 1);
                kelem.push_back(k2[i]);
            }
        base += (n + 1) * (n + 1);
    }
    m.finalize();
    parameter K2;
    K2.setelementvalues(kelem);
    field u(m, "h1", "u");
    u.setorder(1, order);
    p->fields = {u};
    p->params = {K2};
    p->form = std::make_shared<formulation>(m);
    using namespace sl;
    *p->form += integral(1, grad(dof(u)) * grad(tf(u)) + expression(K2) * dof(u) * tf(u));
    std::ostringstream d;
    d << ninst << " quad instances Q" << order << " (" << m.count(QUAD) << " quads), K + k^2 M";
    p->describe = d.str();
    return p;
}

// Structured unit-square / unit-cube grids of one element type.
void grid_mesh(Mesh& m, const std::string& elem, int n, double jit, Rng& rng, int volreg, int bndreg) {
    bool is3d = elem == "tet" || elem == "hex" || elem == "prism";
    int n1 = n + 1;
    double h = 1.0 / n;
    std::vector<double> xyz;
    int nk = is3d ? n : 0;
    for (int k = 0; k <= nk; k++)
        for (int j = 0; j <= n; j++)
            for (int i = 0; i <= n; i++) {
                double c[3] = {i * h, j * h, is3d ? k * h : 0.0};
                bool interior = i > 0 && i < n && j > 0 && j < n && (!is3d || (k > 0 && k < n));
                if (interior && jit > 0)
                    for (int d = 0; d < (is3d ? 3 : 2); d++) c[d] += rng.uniform(-jit * h, jit * h);
                xyz.insert(xyz.end(), {c[0], c[1], c[2]});
            }
    m.set_nodes(xyz);
    auto v2 = [&](int i, int j) { return i + n1 * j; };
    auto v3 = [&](int i, int j, int k) { return i + n1 * (j + n1 * k); };
    if (!is3d) {
        if (bndreg >= 0) square_boundary(m, n, 0, bndreg);
        if (elem == "tri") square_tris(m, n, 0, volreg);
        else
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) m.add_element(QUAD, {v2(i, j), v2(i + 1, j), v2(i + 1, j + 1), v2(i, j + 1)}, volreg);
        return;
    }
    if (elem == "hex") {
        if (bndreg >= 0) {
            // the six faces of the cube as quadrangles
            for (int a = 0; a < n; a++)
                for (int b = 0; b < n; b++) {
                    m.add_element(QUAD, {v3(0, a, b), v3(0, a, b + 1), v3(0, a + 1, b + 1), v3(0, a + 1, b)}, bndreg);
                    m.add_element(QUAD, {v3(n, a, b), v3(n, a + 1, b), v3(n, a + 1, b + 1), v3(n, a, b + 1)}, bndreg);
                    m.add_element(QUAD, {v3(a, 0, b), v3(a + 1, 0, b), v3(a + 1, 0, b + 1), v3(a, 0, b + 1)}, bndreg);
                    m.add_element(QUAD, {v3(a, n, b), v3(a, n, b + 1), v3(a + 1, n, b + 1), v3(a + 1, n, b)}, bndreg);
                    m.add_element(QUAD, {v3(a, b, 0), v3(a, b + 1, 0), v3(a + 1, b + 1, 0), v3(a + 1, b, 0)}, bndreg);
                    m.add_element(QUAD, {v3(a, b, n), v3(a + 1, b, n), v3(a + 1, b + 1, n), v3(a, b + 1, n)}, bndreg);
                }
        }
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++)
                    m.add_element(HEX, {v3(i, j, k), v3(i + 1, j, k), v3(i + 1, j + 1, k), v3(i, j + 1, k), v3(i, j, k + 1),
                                        v3(i + 1, j, k + 1), v3(i + 1, j + 1, k + 1), v3(i, j + 1, k + 1)},
                                  volreg);
        return;
    }
    if (elem == "prism") {
        if (bndreg >= 0) {
            // the four side faces of the cube as quadrangles (prism side faces)
            for (int a = 0; a < n; a++)
                for (int b = 0; b < n; b++) {
                    m.add_element(QUAD, {v3(0, a, b), v3(0, a, b + 1), v3(0, a + 1, b + 1), v3(0, a + 1, b)}, bndreg);
                    m.add_element(QUAD, {v3(n, a, b), v3(n, a + 1, b), v3(n, a + 1, b + 1), v3(n, a, b + 1)}, bndreg);
                    m.add_element(QUAD, {v3(a, 0, b), v3(a + 1, 0, b), v3(a + 1, 0, b + 1), v3(a, 0, b + 1)}, bndreg);
                    m.add_element(QUAD, {v3(a, n, b), v3(a, n, b + 1), v3(a + 1, n, b + 1), v3(a + 1, n, b)}, bndreg);
                }
        }
        // every cube cut along its (i, j) -> (i+1, j+1) diagonal into two
        // prisms, bottom triangle counter-clockwise seen from +z
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    m.add_element(PRISM, {v3(i, j, k), v3(i + 1, j, k), v3(i + 1, j + 1, k), v3(i, j, k + 1), v3(i + 1, j, k + 1),
                                          v3(i + 1, j + 1, k + 1)},
                                  volreg);
                    m.add_element(PRISM, {v3(i, j, k), v3(i + 1, j + 1, k), v3(i, j + 1, k), v3(i, j, k + 1), v3(i + 1, j + 1, k + 1),
                                          v3(i, j + 1, k + 1)},
                                  volreg);
                }
        return;
    }
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++)
                for (auto& pm : perms) {
                    int c[3] = {i, j, k};
                    std::vector<int> tet = {v3(c[0], c[1], c[2])};
                    for (int a = 0; a < 2; a++) {
                        c[pm[a]]++;
                        tet.push_back(v3(c[0], c[1], c[2]));
                    }
                    tet.push_back(v3(i + 1, j + 1, k + 1));
                    m.add_element(TET, tet, volreg);
                }
}

// L2 projection of a polynomial of degree `deg` (dof*tf - f*tf): with
// deg <= order the projection reproduces f exactly, which pins bases,
// orientations, numbering and assembly together (the validation of
// examples/high-order-interpolations-2d/main.cpp:68 in exact form).
// pmax > order: per-element orders uniform in [order, pmax] (one physical
// region per order, set in descending order so that interfaces take the
// lower order, as in the hp workload); f of degree <= order stays exact.
std::shared_ptr<Problem> projection(const std::map<std::string, std::string>& s) {
    auto p = std::make_shared<Problem>();
    std::string elem = s.count("elem") ? s.at("elem") : "tri";
    int n = geti(s, "n", 4), order = geti(s, "order", 3), deg = geti(s, "deg", order);
    const int pmax = geti(s, "pmax", order);
    double jit = getd(s, "jitter", (elem == "tri" || elem == "tet") ? 0.2 : 0.0);
    Rng rng((uint64_t)geti(s, "seed", 7));
    p->mesh = std::make_shared<Mesh>();
    Mesh& m = *p->mesh;
    grid_mesh(m, elem, n, jit, rng, 1, -1);
    const int vtype = elem == "tri" ? TRI : elem == "quad" ? QUAD : elem == "tet" ? TET : elem == "hex" ? HEX : PRISM;
    if (pmax > order)
        for (int e = 0; e < m.count(vtype); e++) m.add_to_physreg(vtype, e, 100 + rng.uniform_int(order, pmax));
    m.finalize();
    field u(m, "h1", "u");
    u.setorder(1, order);
    for (int q = pmax; q > order; q--)
        if (m.has_physreg(100 + q)) u.setorder(100 + q, q);
    if (pmax > order && m.has_physreg(100 + order)) u.setorder(100 + order, order);
    p->fields = {u};
    p->form = std::make_shared<formulation>(m);
    using namespace sl;
    bool is3d = elem == "tet" || elem == "hex" || elem == "prism";
    expression f(0.0);
    bool first = true;
    for (int a = 0; a <= deg; a++)
        for (int b = 0; a + b <= deg; b++)
            for (int c = 0; a + b + c <= deg && (is3d || c == 0); c++) {
                expression mono(1.0 / (1.0 + a + 2.0 * b + 3.0 * c));
                for (int k = 0; k < a; k++) mono = mono * x();
                for (int k = 0; k < b; k++) mono = mono * y();
                for (int k = 0; k < c; k++) mono = mono * z();
                f = first ? mono : f + mono;
                first = false;
            }
    *p->form += integral(1, dof(u) * tf(u) - f * tf(u));
    p->describe = "L2 projection of a degree-" + std::to_string(deg) + " polynomial, " + elem + " order " + std::to_string(order);
    return p;
}

// H(curl) field E of the given order on a unit square / cube grid (nodes
// randomly renumbered with shuffle=1, so that edge and face orientations
// vary), with a vector polynomial f of degree `deg` (third component 0 in 2D):
//   term=mass     integral(dof(E) tf(E) - f tf(E)): the L2 projection, exact
//                 when f lies in the discrete space (degree <= order on
//                 simplices and affine quadrangles / hexahedra, order >= 1);
//   term=curl     integral(curl(dof(E)) curl(tf(E))) + integral(-f tf(E)),
//                 the same DOF numbering as term=mass;
//   term=maxwell  integral(-curl(dof(E)) curl(tf(E)) + k2 dof(E) tf(E) - f tf(E))
//                 (the waveguide examples' formulation with a source).
std::shared_ptr<Problem> hcurlprojection(const std::map<std::string, std::string>& s) {
    auto p = std::make_shared<Problem>();
    const std::string elem = s.count("elem") ? s.at("elem") : "tri";
    const std::string term = s.count("term") ? s.at("term") : "mass";
    const int n = geti(s, "n", 3), order = geti(s, "order", 1), deg = geti(s, "deg", std::max(order, 0));
    const double jit = getd(s, "jitter", (elem == "tri" || elem == "tet") ? 0.2 : 0.0);
    Rng rng((uint64_t)geti(s, "seed", 11));
    p->mesh = std::make_shared<Mesh>();
    Mesh& m = *p->mesh;
    grid_mesh(m, elem, n, jit, rng, 1, -1);
    if (geti(s, "shuffle", 1)) {
        std::vector<int> perm(m.count(0));
        std::iota(perm.begin(), perm.end(), 0);
        for (int i = (int)perm.size() - 1; i > 0; i--) std::swap(perm[i], perm[rng.uniform_int(0, i)]);
        m.permute_nodes(perm);
    }
    m.finalize();
    field E(m, "hcurl", "E");
    E.setorder(1, order);
    p->fields = {E};
    p->form = std::make_shared<formulation>(m);
    using namespace sl;
    const bool is3d = elem == "tet" || elem == "hex";
    std::vector<expression> fc;
    for (int comp = 0; comp < 3; comp++) {
        expression f(0.0);
        bool first = true;
        for (int a = 0; a <= deg && (is3d || comp < 2); a++)
            for (int b = 0; a + b <= deg; b++)
                for (int c = 0; a + b + c <= deg && (is3d || c == 0); c++) {
                    expression mono(((comp + a + b + c) % 2 ? -1.0 : 1.0) / (1.0 + a + 2.0 * b + 3.0 * c + 5.0 * comp));
                    for (int k = 0; k < a; k++) mono = mono * x();
                    for (int k = 0; k < b; k++) mono = mono * y();
                    for (int k = 0; k < c; k++) mono = mono * z();
                    f = first ? mono : f + mono;
                    first = false;
                }
        fc.push_back(f);
    }
    const expression f(3, 1, fc);
    const int delta = geti(s, "delta", 0);
    if (term == "mass") {
        *p->form += integral(1, dof(E) * tf(E) - f * tf(E), delta);
    } else if (term == "curl") {
        *p->form += integral(1, curl(dof(E)) * curl(tf(E)), delta);
        *p->form += integral(1, -f * tf(E), delta);
    } else {
        SLX_CHECK(term == "maxwell", "term is mass, curl or maxwell");
        const expression k2(getd(s, "k2", 25.0));
        *p->form += integral(1, -curl(dof(E)) * curl(tf(E)) + k2 * dof(E) * tf(E) - f * tf(E), delta);
    }
    p->describe = "hcurl " + term + " on " + elem + " order " + std::to_string(order) + ", f of degree " + std::to_string(deg);
    return p;
}

// Config: 3D hp-Poisson on hexahedra, per-element order uniform in
// [pmin, pmax] (one physical region per order, orders set descendingly so
// interfaces take the lower order, rawfield.cpp:670-676), coefficient
// lognormal(0, sigma) per element, Gauss order 2p+1 (delta -1), u = 0 on the
// cube boundary, f = 1.
std::shared_ptr<Problem> hppoisson3d(const std::map<std::string, std::string>& s) {
    auto p = std::make_shared<Problem>();
    int n = geti(s, "n", 48), pmin = geti(s, "pmin", 1), pmax = geti(s, "pmax", 5);
    double jit = getd(s, "jitter", 0.1), sigma = getd(s, "sigma", 1.0);
    Rng rng((uint64_t)geti(s, "seed", 2));
    p->mesh = std::make_shared<Mesh>();
    Mesh& m = *p->mesh;
    grid_mesh(m, "hex", n, jit, rng, 1, 2);
    int nh = m.count(HEX);
    std::vector<int> order(nh);
    for (int e = 0; e < nh; e++) {
        order[e] = rng.uniform_int(pmin, pmax);
        m.add_to_physreg(HEX, e, 100 + order[e]);
    }
    m.finalize();
    parameter c;
    std::vector<double> cv(nh);
    for (int e = 0; e < nh; e++) cv[e] = std::exp(sigma * rng.normal());
    c.setelementvalues(cv);
    field u(m, "h1", "u");
    for (int q = pmax; q >= pmin; q--)
        if (m.has_physreg(100 + q)) u.setorder(100 + q, q);
    u.setconstraint(2);
    p->fields = {u};
    p->params = {c};
    p->form = std::make_shared<formulation>(m);
    using namespace sl;
    *p->form += integral(1, expression(c) * grad(dof(u)) * grad(tf(u)), -1);
    *p->form += integral(1, -expression(1.0) * tf(u), -1);
    std::ostringstream d;
    d << "3D hp-Poisson, " << n << "^3 hexahedra, orders " << pmin << ".." << pmax << ", jitter " << jit;
    p->describe = d.str();
    return p;
}

// 3D Poisson with a mass term on prisms (n^3 cubes, two prisms each,
// jittered interior nodes, field order `order`): integral(c grad(dof(u))
// grad(tf(u)) + dof(u) tf(u)) and -f tf(u), u = 0 on the four side faces
// (closed=1 adds the bottom and top; mass=0 drops the mass term).  The prism
// H1 functions follow h1prism.cpp; the Gauss rule gpprism.cpp.
std::shared_ptr<Problem> prism3d(const std::map<std::string, std::string>& s) {
    auto p = std::make_shared<Problem>();
    int n = geti(s, "n", 32);
    double jit = getd(s, "jitter", 0.1), sigma = getd(s, "sigma", 1.0);
    Rng rng((uint64_t)geti(s, "seed", 3));
    p->mesh = std::make_shared<Mesh>();
    Mesh& m = *p->mesh;
    grid_mesh(m, "prism", n, jit, rng, 1, 2);
    if (geti(s, "closed", 0)) {
        // closed=1: the bottom and top triangles join the Dirichlet region too
        const int n1 = n + 1;
        auto v3 = [&](int i, int j, int k) { return i + n1 * (j + n1 * k); };
        for (int k = 0; k <= n; k += n)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    m.add_element(TRI, {v3(i, j, k), v3(i + 1, j, k), v3(i + 1, j + 1, k)}, 2);
                    m.add_element(TRI, {v3(i, j, k), v3(i + 1, j + 1, k), v3(i, j + 1, k)}, 2);
                }
    }
    m.finalize();
    int np = m.count(PRISM);
    parameter c;
    std::vector<double> cv(np);
    for (int e = 0; e < np; e++) cv[e] = std::exp(sigma * rng.normal());
    c.setelementvalues(cv);
    field u(m, "h1", "u");
    u.setorder(1, geti(s, "order", 1));
    u.setconstraint(2);
    p->fields = {u};
    p->params = {c};
    p->form = std::make_shared<formulation>(m);
    using namespace sl;
    *p->form += integral(1, expression(c) * grad(dof(u)) * grad(tf(u)), geti(s, "delta", 0));
    if (geti(s, "mass", 1)) *p->form += integral(1, dof(u) * tf(u), geti(s, "delta", 0));
    *p->form += integral(1, -expression(getd(s, "f", 1.0)) * tf(u), geti(s, "delta", 0));
    std::ostringstream d;
    d << "3D Poisson + mass, " << n << "^3 x 2 prisms, order " << geti(s, "order", 1) << ", jitter " << jit;
    p->describe = d.str();
    return p;
}

std::shared_ptr<Problem> from_msh(const std::map<std::string, std::string>& s) {
    auto p = std::make_shared<Problem>();
    auto it = s.find("file");
    SLX_CHECK(it != s.end(), "msh problem needs file=<path>");
    p->mesh = std::make_shared<Mesh>(Mesh::read_msh(it->second));
    Mesh& m = *p->mesh;
    int vol = geti(s, "vol", 1), order = geti(s, "order", 1);
    field u(m, "h1", "u");
    u.setorder(vol, order);
    if (s.count("dirichlet")) u.setconstraint(geti(s, "dirichlet", 2));
    p->fields = {u};
    p->form = std::make_shared<formulation>(m);
    using namespace sl;
    std::string integrand = s.count("integrand") ? s.at("integrand") : "poisson";
    if (integrand == "mass")
        *p->form += integral(vol, dof(u) * tf(u), geti(s, "delta", 0));
    else
        *p->form += integral(vol, grad(dof(u)) * grad(tf(u)), geti(s, "delta", 0));
    *p->form += integral(vol, -expression(getd(s, "f", 1.0)) * tf(u), geti(s, "delta", 0));
    p->describe = "mesh file " + it->second;
    return p;
}

// examples/elasticity-membrane-3d/main.cpp:10-40 on its mesh file (a disk of
// 27-node curved hexahedra): h1xyz displacement of order 3 on volume 1,
// clamped on surface 2, E = 150e9, nu = 0.3, volume force (0, 0, -10).
std::shared_ptr<Problem> membrane3d(const std::map<std::string, std::string>& s) {
    auto p = std::make_shared<Problem>();
    auto it = s.find("file");
    SLX_CHECK(it != s.end(), "membrane3d needs file=<path>");
    p->mesh = std::make_shared<Mesh>(Mesh::read_msh(it->second));
    Mesh& m = *p->mesh;
    const int vol = 1, sur = 2;
    field u(m, "h1xyz", "u");
    u.setorder(vol, geti(s, "order", 3));
    u.setconstraint(sur);
    p->fields = {u};
    p->form = std::make_shared<formulation>(m);
    using namespace sl;
    const expression E(getd(s, "E", 150e9)), nu(getd(s, "nu", 0.3));
    *p->form += integral(vol, predefinedelasticity(dof(u), tf(u), E, nu));
    if (geti(s, "measure", 0)) {
        // the validation integral compz(u).integrate(vol, u, 5) (main.cpp:50) as a
        // vector: rhs[i] = integral of the z component of form function i at
        // integration order 5 (delta -3 from the default 3 + 3 + 2), so that the
        // integral of u_z is rhs . x; the same contributions' fields as the
        // problem itself, hence the same DOF numbering
        expression ez(3, 1, {expression(0.0), expression(0.0), expression(-1.0)});
        *p->form += integral(vol, ez * tf(u), -3);
        p->describe = "elasticity membrane 3D: integral of u_z at order 5";
        return p;
    }
    expression f(3, 1, {expression(0.0), expression(0.0), expression(-10.0)});
    *p->form += integral(vol, f * tf(u));
    p->describe = "elasticity membrane 3D (curved hexahedra), mesh file " + it->second;
    return p;
}

// examples/electrostatics-1d/main.cpp:11-36 on its mesh file: P1 potential
// on line region 1, 10 V on point region 2 and 2 V on point region 3,
// integral(line, -epsilon grad(dof(v)) grad(tf(v))).
std::shared_ptr<Problem> electrostatics1d(const std::map<std::string, std::string>& s) {
    auto p = std::make_shared<Problem>();
    auto it = s.find("file");
    SLX_CHECK(it != s.end(), "electrostatics1d needs file=<path>");
    p->mesh = std::make_shared<Mesh>(Mesh::read_msh(it->second));
    Mesh& m = *p->mesh;
    const int line = 1, left = 2, right = 3;
    field v(m, "h1", "v");
    v.setorder(line, geti(s, "order", 1));
    v.setconstraint(left, 10.0);
    v.setconstraint(right, 2.0);
    p->fields = {v};
    p->form = std::make_shared<formulation>(m);
    using namespace sl;
    const double epsilon = 8.854e-12;
    *p->form += integral(line, -expression(epsilon) * grad(dof(v)) * grad(tf(v)));
    p->describe = "electrostatics 1D, mesh file " + it->second;
    return p;
}

// examples/electromagnetic-waveguide-2d/main.cpp:13-40 on its mesh file: hcurl
// electric field of order 2 on the whole domain (region 3), tangential E = 0
// on the skin (region 2), E = sin(y / 0.1 pi) (0, 1, 0) projected on the
// left edge (region 1), integral(-curl(dof(E)) curl(tf(E)) + k^2 dof(E) tf(E))
// with k = 2 pi f / c, f = 0.9 GHz, c = 3e8 and the example's pi = 3.14159.
std::shared_ptr<Problem> waveguide2d(const std::map<std::string, std::string>& s) {
    auto p = std::make_shared<Problem>();
    auto it = s.find("file");
    SLX_CHECK(it != s.end(), "waveguide2d needs file=<path>");
    p->mesh = std::make_shared<Mesh>(Mesh::read_msh(it->second));
    Mesh& m = *p->mesh;
    const int left = 1, skin = 2, wholedomain = 3;
    field E(m, "hcurl", "E");
    E.setorder(wholedomain, geti(s, "order", 2));
    const double freq = getd(s, "freq", 0.9e9), c = 3e8, pi = 3.14159, k = 2 * pi * freq / c;
    using namespace sl;
    E.setconstraint(skin);
    E.setconstraint(left, sl::sin(y() / expression(0.1) * expression(pi)) * expression(3, 1, {expression(0.0), expression(1.0), expression(0.0)}));
    p->fields = {E};
    p->form = std::make_shared<formulation>(m);
    *p->form += integral(wholedomain, -curl(dof(E)) * curl(tf(E)) + expression(k * k) * dof(E) * tf(E));
    p->describe = "electromagnetic waveguide 2D (hcurl), mesh file " + it->second;
    return p;
}

// examples/high-order-interpolations-2d/main.cpp:17-48 on its mesh file:
// L2 projection of sin(10x) sin(13y) onto h1 of the given order on face 1,
// integration order +20.
std::shared_ptr<Problem> highorderprojection2d(const std::map<std::string, std::string>& s) {
    auto p = std::make_shared<Problem>();
    auto it = s.find("file");
    SLX_CHECK(it != s.end(), "hoprojection2d needs file=<path>");
    p->mesh = std::make_shared<Mesh>(Mesh::read_msh(it->second));
    Mesh& m = *p->mesh;
    const int face = 1;
    field v(m, "h1", "v");
    v.setorder(face, geti(s, "order", 5));
    p->fields = {v};
    p->form = std::make_shared<formulation>(m);
    using namespace sl;
    expression toproject = sl::sin(expression(10.0) * x()) * sl::sin(expression(13.0) * y());
    *p->form += integral(face, dof(v) * tf(v) - toproject * tf(v), geti(s, "delta", 20));
    p->describe = "high-order projection 2D, mesh file " + it->second;
    return p;
}

}  // namespace

std::shared_ptr<Problem> make_problem(const std::string& kind, const std::string& spec, void* hip_stream) {
    auto s = parse_spec(spec);
    std::shared_ptr<Problem> p;
    struct StreamScope {
        explicit StreamScope(void* st) { set_mesh_pipeline_stream(st); }
        ~StreamScope() { set_mesh_pipeline_stream(nullptr); }
    } scope(s.count("meshpipe") && s.at("meshpipe") == "host" ? nullptr : hip_stream);
    if (kind == "poisson2d") p = poisson2d(s);
    else if (kind == "elasticity3d") p = elasticity3d(s);
    else if (kind == "quadbatch") p = quadbatch(s);
    else if (kind == "projection") p = projection(s);
    else if (kind == "hcurlprojection") p = hcurlprojection(s);
    else if (kind == "waveguide2d") p = waveguide2d(s);
    else if (kind == "hppoisson3d") p = hppoisson3d(s);
    else if (kind == "prism3d") p = prism3d(s);
    else if (kind == "msh") p = from_msh(s);
    else if (kind == "electrostatics1d") p = electrostatics1d(s);
    else if (kind == "membrane3d") p = membrane3d(s);
    else if (kind == "newton2d") p = newton2d(s);
    else if (kind == "hoprojection2d") p = highorderprojection2d(s);
    else if (kind == "convdiff2d") p = convdiff2d(s);
    else fail("unknown problem kind '" + kind + "'");
    p->kind = kind;
    p->spec = s;
    p->form->set_orientation_folding(geti(s, "fold", 1) != 0);
    p->plan = p->form->buildplan();
    return p;
}

}  // namespace slx
