// Fields, parameters, DOF manager and the formulation front end (see fem.h).
#include "fem.h"
#include <algorithm>
#include <cmath>
#include <functional>
#include <map>
#include <numeric>
#include <tuple>
#include <unordered_map>
#include <cstring>
#include "basis.h"
#include "gauss.h"

namespace slx {

// ---------------------------------------------------------------- fields ----

// Form functions of field f carried by one sub-entity of type t (its own
// dimension) at the given order, and on a whole element of type t.
static int entity_nff(const RawField& f, int t, int order) {
    if (f.hcurl()) return t == POINT ? 0 : hcurl_count(t, order, topo::dimension(t), 0);
    return h1_count(t, order, topo::dimension(t));
}
static int element_nff(const RawField& f, int t, int order) { return f.hcurl() ? hcurl_count(t, order) : h1_count(t, order); }
static std::vector<FormFunction> element_ffs(const RawField& f, int t, int order) {
    return f.hcurl() ? hcurl_iterate(t, order) : h1_iterate(t, order);
}

RawField::RawField(const Mesh* m, const std::string& type, const std::string& name) : mesh_(m), type_(type), name_(name) {
    orders_.assign(m->count_disjregs(), 1);
    constrained_.assign(m->count_disjregs(), 0);
    cvalue_.assign(m->count_disjregs(), 0.0);
    cexpr_.assign(m->count_disjregs(), nullptr);
    ceio_.assign(m->count_disjregs(), 0);
}

void RawField::setorder(int physreg, int order) {
    // hierarchicalformfunction::getminorder: 0 for hcurl, 1 for h1
    SLX_CHECK(order >= (hcurl() ? 0 : 1), "cannot use interpolation order " << order << " for shape function " << type_);
    for (auto& s : subfields_) s->setorder(physreg, order);
    if (!subfields_.empty()) return;
    int regdim = mesh_->physreg_dim(physreg);
    SLX_CHECK(regdim == mesh_->dimension(), "cannot set the interpolation order on a " << regdim << "D region in a "
                                                                                        << mesh_->dimension() << "D problem");
    for (int dr : mesh_->physreg_disjregs(physreg, -1)) orders_[dr] = order;
}

void RawField::setconstraint(int physreg, double value) {
    SLX_CHECK(!hcurl() || value == 0.0, "an hcurl field is constrained with a 3x1 expression (only 0 is provided)");
    for (auto& s : subfields_) s->setconstraint(physreg, value);
    if (!subfields_.empty()) return;
    for (int dr : mesh_->physreg_disjregs(physreg, -1)) {
        constrained_[dr] = 1;
        cvalue_[dr] = value;
        cexpr_[dr] = nullptr;
    }
}

void RawField::setconstraint(int physreg, const expression& value, int extraintegrationdegree) {
    SLX_CHECK(hcurl(), "expression constraints are provided for hcurl fields (h1 fields take a constant)");
    SLX_CHECK(value.countrows() == 3 && value.countcolumns() == 1, "the rawfield must be constrained using a 3x1 expression");
    for (int r = 0; r < 3; r++)
        SLX_CHECK(!value.op(r, 0)->isdofincluded() && !value.op(r, 0)->istfincluded(), "a constraint cannot contain dof() or tf()");
    auto e = std::make_shared<expression>(value);
    for (int dr : mesh_->physreg_disjregs(physreg, -1)) {
        constrained_[dr] = 1;
        cvalue_[dr] = 0.0;
        cexpr_[dr] = e;
        ceio_[dr] = extraintegrationdegree;
    }
}

void RawField::build_data_numbering() {
    // The field's own coefficient numbering: every disjoint region in order,
    // form functions in order, elements in order (like addtostructure on all
    // disjoint regions, dofmanager.cpp:35-123).
    int nd = mesh_->count_disjregs();
    data_rangebegin.assign(nd, {});
    int64_t n = 0;
    for (int dr = 0; dr < nd; dr++) {
        int nff = entity_nff(*this, mesh_->disjreg_type(dr), orders_[dr]);
        for (int ff = 0; ff < nff; ff++) {
            data_rangebegin[dr].push_back((int)n);
            n += mesh_->disjreg_count(dr);
        }
    }
    data_size_ = n;
    data.assign(n, 0.0);
}

field::field(const Mesh& m, const std::string& type, const std::string& name) {
    int ncomp = 1;
    std::string base = type;
    if (type.size() > 2 && type.compare(type.size() - 3, 3, "xyz") == 0) { ncomp = 3; base = type.substr(0, type.size() - 3); }
    else if (type.size() > 2 && type.compare(type.size() - 2, 2, "xy") == 0) { ncomp = 2; base = type.substr(0, type.size() - 2); }
    SLX_CHECK(base == "h1" || (base == "hcurl" && ncomp == 1),
              "field types are h1, h1xy, h1xyz and hcurl (got '" << type << "')");
    raw_ = std::make_shared<RawField>(&m, base, name);
    raw_->setself(raw_);
    if (ncomp > 1) {
        const char* comp = "xyz";
        for (int i = 0; i < ncomp; i++) {
            auto s = std::make_shared<RawField>(&m, base, name + "." + comp[i]);
            s->setself(s);
            raw_->addsub(s);
        }
    }
}

void field::setorder(int physreg, int order) { raw_->setorder(physreg, order); }
void field::setconstraint(int physreg, double value) { raw_->setconstraint(physreg, value); }
void field::setconstraint(int physreg, const expression& value, int extraintegrationdegree) {
    raw_->setconstraint(physreg, value, extraintegrationdegree);
}

static OpP field_op(const std::shared_ptr<RawField>& f) {
    auto o = std::make_shared<Op>();
    o->kind = OpKind::Field;
    o->field = f;
    return o;
}

field::operator expression() const {
    if (raw_->hcurl()) {
        // reference components v_r mapped by J^-T (expression.cpp:30-35 with
        // expression::invjac, expression.cpp:2031-2050): row k = sum_r
        // invjac(k, r) v_r over the mesh dimension; the third row of a 2D
        // mesh multiplies the (identically zero) third reference component
        const int md = raw_->meshdim();
        SLX_CHECK(md >= 2, "hcurl fields need a 2D or 3D mesh");
        std::vector<expression> rows;
        for (int k = 0; k < 3; k++) {
            if (k >= md) {
                rows.push_back(expression(0.0));
                continue;
            }
            expression e(0.0);
            for (int r = 0; r < md; r++) {
                OpP v = field_op(raw_);
                v->hc = 1;
                v->kietaphideriv = r + 1;
                auto ij = std::make_shared<Op>();
                ij->kind = OpKind::InvJac;
                ij->row = k;
                ij->col = r;
                expression t = expression(ij) * expression(v);
                e = r == 0 ? t : e + t;
            }
            rows.push_back(e);
        }
        expression out(3, 1, rows);
        out.set_hcurl(raw_, OpKind::Field, -1);
        return out;
    }
    int n = raw_->ncomponents();
    if (n == 1) return expression(field_op(raw_));
    std::vector<expression> e;
    for (int i = 0; i < n; i++) e.push_back(expression(field_op(raw_->component(i))));
    return expression(n, 1, e);
}

expression field::comp(int i) const {
    if (raw_->hcurl()) return expression(*this).at(i, 0);
    return expression(field_op(raw_->component(i)));
}

parameter::operator expression() const {
    auto o = std::make_shared<Op>();
    o->kind = OpKind::Parameter;
    o->param = raw_;
    return expression(o);
}

// ------------------------------------------------------------ sl functions ----

namespace sl {
expression dof(const expression& e, int physreg) { return e.dof(physreg); }
expression tf(const expression& e, int physreg) { return e.tf(physreg); }
expression dx(const expression& e) { return e.spacederivative(1); }
expression dy(const expression& e) { return e.spacederivative(2); }
expression dz(const expression& e) { return e.spacederivative(3); }

static int meshdim_of(const expression& e) {
    for (int r = 0; r < e.countrows(); r++)
        for (int c = 0; c < e.countcolumns(); c++)
            if (e.op(r, c)->field) return e.op(r, c)->field->meshdim();
    fail("cannot take the gradient of an expression without fields");
}

expression grad(const expression& in) {
    // sl::grad (sl.cpp:1088-1128): entry (comp, i) = d/dx_i of component comp.
    SLX_CHECK(in.countcolumns() == 1 && in.countrows() <= 3, "can only take the gradient of a scalar or an up to length 3 column vector");
    int pd = meshdim_of(in);
    std::vector<expression> ex;
    for (int comp = 0; comp < in.countrows(); comp++)
        for (int i = 0; i < pd; i++) ex.push_back(in.spacederivative(i + 1).at(comp, 0));
    expression out(in.countrows(), pd, ex);
    if (in.countrows() == 1) out = out.transpose();
    return out;
}

expression curl(const expression& in) {
    if (const std::shared_ptr<RawField>& f = in.hcurl_field()) {
        // curl in the reference element, brought back with J^T / det J
        // (sl.cpp:1215-1222); a 2D element's curl is along z only
        const int md = f->meshdim();
        auto leaf = [&](int r) {
            auto o = std::make_shared<Op>();
            o->kind = in.hcurl_kind();
            o->field = f;
            o->physreg = in.hcurl_physreg();
            o->hc = 2;
            o->kietaphideriv = r + 1;
            return expression(o);
        };
        auto detj = std::make_shared<Op>();
        detj->kind = OpKind::DetJac;
        const expression invdet = expression(1.0) / expression(detj);
        std::vector<expression> rows;
        for (int k = 0; k < 3; k++) {
            if (md == 2) {
                rows.push_back(k == 2 ? leaf(2) * invdet : expression(0.0));
                continue;
            }
            expression e(0.0);
            for (int r = 0; r < 3; r++) {
                auto j = std::make_shared<Op>();
                j->kind = OpKind::Jac;
                j->row = r;
                j->col = k;
                expression t = expression(j) * leaf(r);
                e = r == 0 ? t : e + t;
            }
            rows.push_back(e * invdet);
        }
        return expression(3, 1, rows);
    }
    SLX_CHECK(in.countcolumns() == 1 && in.countrows() <= 3, "can only take the curl of an up to length 3 column vector");
    // sl.cpp:1203-1211 (not axisymmetric)
    const expression zero(0.0);
    switch (in.countrows()) {
        case 1: return expression(3, 1, {zero, zero, zero});
        case 2: return expression(3, 1, {zero, zero, dx(in).at(1, 0) - dy(in).at(0, 0)});
    }
    return expression(3, 1, {dy(in).at(2, 0) - dz(in).at(1, 0), dz(in).at(0, 0) - dx(in).at(2, 0), dx(in).at(1, 0) - dy(in).at(0, 0)});
}

expression strain(const expression& input) {
    // sl::strain (sl.cpp:2284-2303)
    expression g = input.countcolumns() == 1 ? grad(input) : input;
    if (input.countrows() == 2) return expression(3, 1, {g.at(0, 0), g.at(1, 1), g.at(1, 0) + g.at(0, 1)});
    SLX_CHECK(input.countrows() == 3, "can only compute the strains of a 2x1 or 3x1 column vector");
    return expression(6, 1, {g.at(0, 0), g.at(1, 1), g.at(2, 2), g.at(2, 1) + g.at(1, 2), g.at(0, 2) + g.at(2, 0), g.at(0, 1) + g.at(1, 0)});
}

expression predefinedelasticity(const expression& dofu, const expression& tfu, const expression& E, const expression& nu) {
    // sl::predefinedelasticity (sl.cpp:2672-2752), 3D isotropic case.
    SLX_CHECK(dofu.countrows() == 3 && tfu.countrows() == 3, "only the 3D isotropic elasticity form is provided");
    expression one(1.0), half(0.5), zero(0.0), two(2.0);
    expression H(6, 6, {one - nu, nu, nu, zero, zero, zero, nu, one - nu, nu, zero, zero, zero, nu, nu, one - nu, zero, zero, zero,
                        zero, zero, zero, half * (one - two * nu), zero, zero, zero, zero, zero, zero, half * (one - two * nu), zero,
                        zero, zero, zero, zero, zero, half * (one - two * nu)});
    expression coef = E / (one + nu) / (one - two * nu);
    coef.reuseit();
    H = coef * H;
    return -(H * strain(dofu)) * strain(tfu);
}

expression pow(const expression& b, const expression& e) { return expression(make_op(OpKind::Power, {b.op(), e.op()})); }
expression abs(const expression& e) { return expression(make_op(OpKind::Abs, {e.op()})); }
expression sqrt(const expression& e) { return pow(e, expression(0.5)); }
expression exp(const expression& e) { return expression(make_op(OpKind::Exp, {e.op()})); }
expression sin(const expression& e) { return expression(make_op(OpKind::Sin, {e.op()})); }
expression cos(const expression& e) { return expression(make_op(OpKind::Cos, {e.op()})); }
expression norm(const expression& e) {
    expression s(0.0);
    bool first = true;
    for (int r = 0; r < e.countrows(); r++)
        for (int c = 0; c < e.countcolumns(); c++) {
            expression t = e.at(r, c) * e.at(r, c);
            s = first ? t : s + t;
            first = false;
        }
    return sqrt(s);
}
static expression coord(int axis) {
    auto o = std::make_shared<Op>();
    o->kind = OpKind::Coord;
    o->row = axis;
    return expression(o);
}
expression x() { return coord(0); }
expression y() { return coord(1); }
expression z() { return coord(2); }
integration integral(int physreg, const expression& e, int delta) { return integration{physreg, e, delta}; }
}  // namespace sl

// ------------------------------------------------------------- dofmanager ----

int DofManager::fieldindex(const RawField* f) const {
    for (size_t i = 0; i < fields_.size(); i++)
        if (fields_[i].get() == f) return (int)i;
    return -1;
}

void DofManager::addtostructure(const std::shared_ptr<RawField>& f, int physreg) {
    addtostructure(f, mesh_->physreg_disjregs(physreg, -1));
}

void DofManager::addtostructure(const std::shared_ptr<RawField>& f, const std::vector<int>& disjregs) {
    // dofmanager::addtostructure (dofmanager.cpp:35-123)
    int fi = fieldindex(f.get());
    if (fi < 0) {
        fi = (int)fields_.size();
        fields_.push_back(f);
        orders_at_add_.push_back(f->orders());
        rangebegin_.emplace_back(mesh_->count_disjregs());
    } else
        SLX_CHECK(f->orders() == orders_at_add_[fi], "field order was changed and does not match dof structure anymore");
    for (int dr : disjregs) {
        int nff = entity_nff(*f, mesh_->disjreg_type(dr), f->order(dr));
        std::vector<int>& rb = rangebegin_[fi][dr];
        int before = (int)rb.size();
        if (before >= nff) continue;
        rb.resize(nff);
        int n = mesh_->disjreg_count(dr);
        for (int ff = before; ff < nff; ff++) {
            rb[ff] = (int)ndofs_;
            ndofs_ += n;
        }
    }
    SLX_CHECK(ndofs_ < ((int64_t)1 << 31), "more than 2^31 DOFs: the reference's int addressing overflows");
}

std::vector<uint8_t> DofManager::isconstrained() const {
    // dofmanager::isconstrained / getdisjregconstrainedindexes (dofmanager.cpp:366-451)
    std::vector<uint8_t> out(ndofs_, 0);
    for (size_t fi = 0; fi < fields_.size(); fi++)
        for (int dr = 0; dr < (int)rangebegin_[fi].size(); dr++) {
            if (rangebegin_[fi][dr].empty() || !fields_[fi]->constrained(dr)) continue;
            int n = mesh_->disjreg_count(dr);
            for (int rb : rangebegin_[fi][dr])
                for (int i = 0; i < n; i++) out[rb + i] = 1;
        }
    return out;
}

// Value of a coefficient expression (no dof, tf or field) at a physical point.
static double eval_at(const OpP& op, const double* xyz) {
    switch (op->kind) {
        case OpKind::Constant: return op->value;
        case OpKind::Coord: return xyz[op->row];
        case OpKind::Sum: {
            double v = 0.0;
            for (const OpP& a : op->args) v += eval_at(a, xyz);
            return v;
        }
        case OpKind::Product: {
            double v = 1.0;
            for (const OpP& a : op->args) v *= eval_at(a, xyz);
            return v;
        }
        case OpKind::Inversion: return 1.0 / eval_at(op->args[0], xyz);
        case OpKind::Power: return std::pow(eval_at(op->args[0], xyz), eval_at(op->args[1], xyz));
        case OpKind::Abs: return std::abs(eval_at(op->args[0], xyz));
        case OpKind::Exp: return std::exp(eval_at(op->args[0], xyz));
        case OpKind::Log: return std::log(eval_at(op->args[0], xyz));
        case OpKind::Sin: return std::sin(eval_at(op->args[0], xyz));
        case OpKind::Cos: return std::cos(eval_at(op->args[0], xyz));
        case OpKind::Tan: return std::tan(eval_at(op->args[0], xyz));
        case OpKind::Sqrt: return std::sqrt(eval_at(op->args[0], xyz));
        default: fail("constraint expressions may hold constants, coordinates and functions of them (got " + op_str(op) + ")");
    }
}

// Dense solve with partial pivoting (the small per-edge projection systems).
static std::vector<double> dense_solve(std::vector<double> A, std::vector<double> b) {
    const int n = (int)b.size();
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::abs(A[(size_t)r * n + c]) > std::abs(A[(size_t)piv * n + c])) piv = r;
        SLX_CHECK(A[(size_t)piv * n + c] != 0.0, "singular constraint projection");
        if (piv != c) {
            for (int k = 0; k < n; k++) std::swap(A[(size_t)c * n + k], A[(size_t)piv * n + k]);
            std::swap(b[c], b[piv]);
        }
        for (int r = c + 1; r < n; r++) {
            const double f = A[(size_t)r * n + c] / A[(size_t)c * n + c];
            for (int k = c; k < n; k++) A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
            b[r] -= f * b[c];
        }
    }
    for (int r = n - 1; r >= 0; r--) {
        double v = b[r];
        for (int k = r + 1; k < n; k++) v -= A[(size_t)r * n + k] * b[k];
        b[r] = v / A[(size_t)r * n + r];
    }
    return b;
}

// hcurl constraint expression on the edges of a 2D mesh: per line element the
// L2 projection of the expression on the element's edge functions
// (rawfield::updateothershapefunctions, rawfield.cpp:164-260, dimension 1:
// block-diagonal projection integral(dof(E) tf(E) - expr tf(E)) at integration
// order 2p + 2 + extra).  On a line of a 2D mesh the Jacobian is completed by
// the unit normal (jacobian.cpp:138-185), so E = v t / |dx/dki| with t the unit
// tangent and det J = |dx/dki|: the mass is int v_i v_j / |dx/dki| dki and the
// load int (expr . t) v_i dki.
static void project_hcurl_edges(const Mesh& m, const RawField& f, int dr, const std::vector<int>& rb, std::vector<double>& out) {
    const expression& ex = *f.constraint_expression(dr);
    const int p = f.order(dr), nff = p + 1;
    SLX_CHECK((int)rb.size() == nff, "constrained edge region without its edge functions");
    const GaussRule rule = gauss_rule(LINE, 2 * p + 2 + f.constraint_integration_degree(dr));
    const int np = rule.count();
    const std::vector<double>& xyz = m.coords();
    std::map<int, std::vector<double>> tabs;  // per total orientation
    for (int e = m.disjreg_begin(dr); e <= m.disjreg_end(dr); e++) {
        const int o = m.total_orientation(LINE, e);
        auto it = tabs.find(o);
        if (it == tabs.end()) {
            std::vector<double> t;
            hcurl_tables(LINE, p, o, HCURL_VALUE, rule.coords, t);
            it = tabs.emplace(o, std::move(t)).first;
        }
        const std::vector<double>& tab = it->second;
        const int n0 = m.subelement(0, LINE, e, 0), n1 = m.subelement(0, LINE, e, 1);
        double a[3], x0[3];
        for (int d = 0; d < 3; d++) {
            x0[d] = xyz[3 * (size_t)n0 + d];
            a[d] = 0.5 * (xyz[3 * (size_t)n1 + d] - x0[d]);  // dx/dki of the order-1 line
        }
        const double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        std::vector<double> M((size_t)nff * nff, 0.0), b(nff, 0.0);
        for (int q = 0; q < np; q++) {
            const double ki = rule.coords[3 * (size_t)q], w = rule.weights[q];
            double x[3];
            for (int d = 0; d < 3; d++) x[d] = x0[d] + (ki + 1.0) * a[d];
            double et = 0.0;
            for (int d = 0; d < 3; d++) et += eval_at(ex.op(d, 0), x) * a[d] / len;
            for (int i = 0; i < nff; i++) {
                const double vi = tab[((size_t)1 * nff + i) * np + q];
                b[i] += w * et * vi;
                for (int j = 0; j < nff; j++) M[(size_t)i * nff + j] += w * vi * tab[((size_t)1 * nff + j) * np + q] / len;
            }
        }
        const std::vector<double> c = dense_solve(M, b);
        for (int i = 0; i < nff; i++) out[(size_t)rb[i] + (e - m.disjreg_begin(dr))] = c[i];
    }
}

std::vector<double> DofManager::constraint_values() const {
    // The reference projects the constraint expression onto the constrained
    // region (rawfield::setdisjregconstraint, rawfield.cpp:1166-1248).  For a
    // constant c the projection is exact in the hierarchical H1 basis: every
    // vertex function (the nodal disjoint regions) gets c, every edge, face
    // and volume function 0.
    std::vector<double> out(ndofs_, 0.0);
    for (size_t fi = 0; fi < fields_.size(); fi++)
        for (int dr = 0; dr < (int)rangebegin_[fi].size(); dr++) {
            if (rangebegin_[fi][dr].empty() || !fields_[fi]->constrained(dr)) continue;
            if (fields_[fi]->hcurl()) {
                if (!fields_[fi]->constraint_expression(dr)) continue;
                const int t = mesh_->disjreg_type(dr);
                if (t == POINT) continue;
                SLX_CHECK(t == LINE && mesh_->dimension() == 2,
                          "hcurl constraints with an expression are provided on line regions of 2D meshes");
                SLX_CHECK(!mesh_->curved(), "hcurl constraint expressions need straight line elements");
                project_hcurl_edges(*mesh_, *fields_[fi], dr, rangebegin_[fi][dr], out);
                continue;
            }
            if (mesh_->disjreg_type(dr) != 0) continue;
            int n = mesh_->disjreg_count(dr);
            int rb = rangebegin_[fi][dr][0];
            for (int i = 0; i < n; i++) out[rb + i] = fields_[fi]->constraint_value(dr);
        }
    return out;
}

void DofManager::setdata(const double* x, int64_t n) const {
    // vec -> field data (sl::setdata, rawfield::setdata through the dof
    // structure): same disjoint region and form function index, element offset
    SLX_CHECK(n == ndofs_, "solution vector has " << n << " entries, the dof structure " << ndofs_);
    for (size_t fi = 0; fi < fields_.size(); fi++) {
        RawField& f = *fields_[fi];
        if (f.data_size() == 0) continue;
        for (int dr = 0; dr < (int)rangebegin_[fi].size(); dr++) {
            const std::vector<int>& rb = rangebegin_[fi][dr];
            const std::vector<int>& db = f.data_rangebegin[dr];
            const int cnt = mesh_->disjreg_count(dr);
            for (size_t ff = 0; ff < rb.size() && ff < db.size(); ff++)
                for (int i = 0; i < cnt; i++) f.data[(size_t)db[ff] + i] = x[(size_t)rb[ff] + i];
        }
    }
}

std::vector<int32_t> DofManager::getaddresses(const std::shared_ptr<RawField>& f, int order, int elemtype,
                                              const std::vector<int>& elems, int fieldphysreg) const {
    // dofmanager::getaddresses (dofmanager.cpp:846-918)
    int fi = fieldindex(f.get());
    SLX_CHECK(fi >= 0, "selected field is not defined in the dof manager");
    std::vector<char> defined(mesh_->count_disjregs(), 0);
    for (int dr : mesh_->physreg_disjregs(fieldphysreg, -1)) defined[dr] = 1;
    std::vector<FormFunction> ffs = element_ffs(*f, elemtype, order);
    int ne = (int)elems.size();
    std::vector<int32_t> out((size_t)ffs.size() * ne);
    for (size_t k = 0; k < ffs.size(); k++) {
        const FormFunction& ff = ffs[k];
        int prevdr = -1, cnt = 0;
        for (int i = 0; i < ne; i++) {
            int sub = mesh_->subelement(ff.assoc_type, elemtype, elems[i], ff.sub);
            int dr = mesh_->disjreg_of(ff.assoc_type, sub);
            if (i == 0 || dr != prevdr) cnt = entity_nff(*f, ff.assoc_type, f->order(dr));
            prevdr = dr;
            int32_t a = -1;
            if (defined[dr] && ff.index < cnt) a = rangebegin_[fi][dr][ff.index] + (sub - mesh_->disjreg_begin(dr));
            out[k * ne + i] = a;
        }
    }
    return out;
}

// ------------------------------------------------------------ formulation ----

formulation::formulation(const Mesh& m) : mesh_(&m), dofm_(std::make_shared<DofManager>(&m)) {}

formulation& formulation::operator+=(const integration& in) {
    // formulation::operator+=(integration) (formulation.cpp:51-149)
    int elementdimension = mesh_->physreg_dim(in.physreg);
    expression e = in.expr;
    SLX_CHECK(e.isscalar(), "expand is only defined for scalar expressions (did you try to define a nonscalar formulation term?)");
    e.op() = op_expand(e.op());
    DofTfSlices s = extract_dof_tf(e.op(), elementdimension);
    int id = nintegrals_++;
    for (size_t sl = 0; sl < s.tfs.size(); sl++) {
        Contribution c;
        c.integral_id = id;
        c.doffield = s.dofs[sl][0] ? s.dofs[sl][0]->field : nullptr;
        c.tffield = s.tfs[sl][0]->field;
        c.doffam = s.dofs[sl][0] ? s.dofs[sl][0]->hc : 0;
        c.tffam = s.tfs[sl][0]->hc;
        c.coeffs = s.coeffs[sl];
        c.dofs = s.dofs[sl];
        c.tfs = s.tfs[sl];
        c.intphysreg = in.physreg;
        c.dofphysreg = (c.doffield && s.dofs[sl][0]->physreg >= 0) ? s.dofs[sl][0]->physreg : in.physreg;
        c.tfphysreg = s.tfs[sl][0]->physreg >= 0 ? s.tfs[sl][0]->physreg : in.physreg;
        c.delta = in.delta;
        if (c.doffield) dofm_->addtostructure(c.doffield, c.dofphysreg);
        dofm_->addtostructure(c.tffield, c.tfphysreg);
        contribs_[c.doffield ? 1 : 0].push_back(c);
    }
    return *this;
}

namespace {

// Coefficient program compiler: CSE by structural signature, left-to-right
// evaluation of sums and products like operation::interpolate
// (opsum.cpp:15-58, opproduct.cpp:4-99), linear-scan register reuse.
struct Compiler {
    std::vector<slx_instr> code;
    std::vector<double>* consts;
    std::unordered_map<std::string, int> memo;       // signature -> value id
    std::vector<std::string> sigs;
    int nvals = 0;
    bool uses_q = false;      // coordinates or fields: varies per Gauss point
    bool uses_geo = false;    // Jacobian entries
    std::vector<std::shared_ptr<RawParameter>>* params;
    std::vector<std::shared_ptr<RawField>>* fields;
    const Mesh* mesh;
    std::vector<int> disjregs;

    int emit(int16_t op, int a, int b, int imm, const std::string& sig) {
        auto it = memo.find(sig);
        if (it != memo.end()) return it->second;
        slx_instr in;
        in.op = op;
        in.dst = (int16_t)nvals;
        in.a = (int16_t)a;
        in.b = (int16_t)b;
        in.imm = imm;
        code.push_back(in);
        memo[sig] = nvals;
        return nvals++;
    }
    int constant(double v) {
        int k = -1;
        for (size_t i = 0; i < consts->size(); i++)
            if (std::memcmp(&(*consts)[i], &v, sizeof(double)) == 0) k = (int)i;
        if (k < 0) { k = (int)consts->size(); consts->push_back(v); }
        std::ostringstream s;
        s << "c" << k;
        return emit(SLX_OP_CONST, 0, 0, k, s.str());
    }
    int compile(const OpP& op) {
        std::ostringstream s;
        switch (op->kind) {
            case OpKind::Constant: return constant(op->value);
            case OpKind::Parameter: {
                auto& p = op->param;
                if (!p->element_values.empty()) {
                    int k = -1;
                    for (size_t i = 0; i < params->size(); i++)
                        if ((*params)[i] == p) k = (int)i;
                    if (k < 0) { k = (int)params->size(); params->push_back(p); }
                    s << "p" << k;
                    return emit(SLX_OP_PARAM, 0, 0, k, s.str());
                }
                // region-wise value: constant on the integration disjoint regions
                // (opparameter::simplify resolves it per disjoint region group)
                double v = 0;
                bool found = false;
                for (auto& kv : p->region_values) {
                    std::vector<int> drs = mesh->physreg_disjregs(kv.first, -1);
                    bool all = true;
                    for (int d : disjregs) all = all && std::find(drs.begin(), drs.end(), d) != drs.end();
                    if (all) { v = kv.second; found = true; }
                }
                SLX_CHECK(found, "parameter is undefined on the integration region");
                return constant(v);
            }
            case OpKind::InvJac:
                uses_geo = true;
                s << "ij" << op->row << op->col;
                return emit(SLX_OP_INVJAC, op->row, op->col, 0, s.str());
            case OpKind::Jac:
                uses_geo = true;
                s << "j" << op->row << op->col;
                return emit(SLX_OP_JAC, op->row, op->col, 0, s.str());
            case OpKind::DetJac:
                uses_geo = true;
                return emit(SLX_OP_DETJ, 0, 0, 0, "dj");
            case OpKind::Coord:
                uses_q = true;
                s << "x" << op->row;
                return emit(SLX_OP_COORD, op->row, 0, 0, s.str());
            case OpKind::Field: {
                uses_q = true;
                int k = -1;
                for (size_t i = 0; i < fields->size(); i++)
                    if ((*fields)[i] == op->field) k = (int)i;
                if (k < 0) { k = (int)fields->size(); fields->push_back(op->field); }
                SLX_CHECK(op->kietaphideriv == 0 && !op->hc,
                          "reference-coordinate derivatives of fields and hcurl field values are not supported in coefficients");
                s << "f" << k << "_" << op->spacederiv;
                return emit(SLX_OP_FIELD, op->spacederiv, 0, k, s.str());
            }
            case OpKind::Sum:
            case OpKind::Product: {
                int r = compile(op->args[0]);
                std::string acc = sigs_of(r);
                for (size_t i = 1; i < op->args.size(); i++) {
                    int b = compile(op->args[i]);
                    std::string sig = std::string(op->kind == OpKind::Sum ? "+(" : "*(") + acc + "," + sigs_of(b) + ")";
                    r = emit(op->kind == OpKind::Sum ? SLX_OP_ADD : SLX_OP_MUL, r, b, 0, sig);
                    acc = sig;
                }
                return r;
            }
            case OpKind::Power: {
                int a = compile(op->args[0]), b = compile(op->args[1]);
                return emit(SLX_OP_POW, a, b, 0, "^(" + sigs_of(a) + "," + sigs_of(b) + ")");
            }
            default: {
                int16_t code_;
                const char* nm;
                switch (op->kind) {
                    case OpKind::Inversion: code_ = SLX_OP_INV; nm = "inv"; break;
                    case OpKind::Abs: code_ = SLX_OP_ABS; nm = "abs"; break;
                    case OpKind::Exp: code_ = SLX_OP_EXP; nm = "exp"; break;
                    case OpKind::Log: code_ = SLX_OP_LOG; nm = "log"; break;
                    case OpKind::Sin: code_ = SLX_OP_SIN; nm = "sin"; break;
                    case OpKind::Cos: code_ = SLX_OP_COS; nm = "cos"; break;
                    case OpKind::Tan: code_ = SLX_OP_TAN; nm = "tan"; break;
                    case OpKind::Sqrt: code_ = SLX_OP_SQRT; nm = "sqrt"; break;
                    default: fail("operation not supported in coefficients: " + op_str(op));
                }
                int a = compile(op->args[0]);
                return emit(code_, a, 0, 0, std::string(nm) + "(" + sigs_of(a) + ")");
            }
        }
    }
    std::string sigs_of(int v) {
        for (auto& kv : memo)
            if (kv.second == v) return kv.first;
        return "?";
    }
};

// Linear-scan register allocation over the SSA program (values die at last use;
// OUT instructions read their operand).
int allocate_registers(std::vector<slx_instr>& code) {
    int n = (int)code.size();
    std::vector<int> lastuse(n, -1);
    auto reads = [](const slx_instr& in, int& na) {
        switch (in.op) {
            case SLX_OP_ADD: case SLX_OP_MUL: case SLX_OP_POW: na = 2; break;
            case SLX_OP_INV: case SLX_OP_ABS: case SLX_OP_EXP: case SLX_OP_LOG: case SLX_OP_SIN:
            case SLX_OP_COS: case SLX_OP_TAN: case SLX_OP_SQRT: case SLX_OP_OUT: na = 1; break;
            default: na = 0;
        }
    };
    // value id v is produced by the instruction whose dst == v (before allocation)
    std::vector<int> producer(n + 1, -1);
    for (int i = 0; i < n; i++)
        if (code[i].op != SLX_OP_OUT) producer[code[i].dst] = i;
    for (int i = 0; i < n; i++) {
        int na;
        reads(code[i], na);
        if (na >= 1) lastuse[code[i].a] = i;
        if (na >= 2) lastuse[code[i].b] = i;
    }
    std::vector<int> reg_of(n + 1, -1);
    std::vector<int> freeregs;
    int nregs = 0;
    for (int i = 0; i < n; i++) {
        slx_instr& in = code[i];
        int na;
        reads(in, na);
        int va = in.a, vb = in.b;
        if (na >= 1) in.a = (int16_t)reg_of[va];
        if (na >= 2) in.b = (int16_t)reg_of[vb];
        // free operands dying here
        if (na >= 1 && lastuse[va] == i) freeregs.push_back(reg_of[va]);
        if (na >= 2 && lastuse[vb] == i && vb != va) freeregs.push_back(reg_of[vb]);
        if (in.op == SLX_OP_OUT) continue;
        int v = in.dst, r;
        if (!freeregs.empty()) { r = freeregs.back(); freeregs.pop_back(); }
        else r = nregs++;
        reg_of[v] = r;
        in.dst = (int16_t)r;
        if (lastuse[v] < 0) freeregs.push_back(r);  // dead value
    }
    return std::max(nregs, 1);
}

// disjointregionselector (disjointregionselector.cpp:4-53)
std::vector<std::vector<int>> select_groups(const Mesh& m, const std::vector<int>& drs, const std::vector<int>& c1,
                                            const std::vector<int>& c2) {
    std::vector<int> idx(drs.size());
    std::iota(idx.begin(), idx.end(), 0);
    auto key = [&](int i) { return std::make_tuple(c1[i], c2[i], m.disjreg_type(drs[i])); };
    std::sort(idx.begin(), idx.end(), [&](int a, int b) {
        if (key(a) != key(b)) return key(a) < key(b);
        return a < b;
    });
    std::vector<std::vector<int>> groups;
    for (size_t k = 0; k < idx.size(); k++) {
        if (k == 0 || key(idx[k]) != key(idx[k - 1])) groups.emplace_back();
        groups.back().push_back(drs[idx[k]]);
    }
    return groups;
}

}  // namespace

std::shared_ptr<PlanStorage> formulation::buildplan() {
    auto ps = std::make_shared<PlanStorage>();
    struct Frag {
        const Contribution* c;
        int m, type, ptf, pdof, intorder, orient;
        bool orientdep;
        bool folded = false;        // all orientations, orientation-0 tables, signed renumbering
        std::vector<int> elems, disjregs;
        std::vector<int> eorient;   // folded: total orientation of every element
    };
    std::vector<Frag> frags;
    // orientation folding maps per (basis, type, order, integration order);
    // basis 0 H1, 1 hcurl
    std::map<std::tuple<int, int, int, int>, H1Fold> folds;
    auto fold_of = [&](int basis, int type, int order, int intorder) -> const H1Fold& {
        auto key = std::make_tuple(basis, type, order, intorder);
        auto it = folds.find(key);
        if (it == folds.end()) {
            const std::vector<double>& pts = gauss_rule(type, intorder).coords;
            it = folds.emplace(key, basis ? hcurl_fold(type, order, pts) : h1_fold(type, order, pts)).first;
        }
        return it->second;
    };
    auto basis_of = [](const std::shared_ptr<RawField>& f) { return f && f->hcurl() ? 1 : 0; };
    auto reads_fields = [](const Contribution& c) {
        std::function<bool(const OpP&)> walk = [&](const OpP& op) {
            if (op->kind == OpKind::Field) return true;
            for (const OpP& a : op->args)
                if (walk(a)) return true;
            return false;
        };
        for (const OpP& co : c.coeffs)
            if (walk(co)) return true;
        return false;
    };
    // formulation::generate (formulation.cpp:211-218) loops rhs first, then K;
    // contribution::generate (contribution.cpp:44-133) builds groups and
    // orientation batches.
    for (int m = 0; m < 2; m++)
        for (const Contribution& c : contribs_[m]) {
            std::vector<int> drs = mesh_->physreg_disjregs(c.intphysreg, -2);
            std::vector<int> tfo(drs.size()), dofo(drs.size(), 0);
            for (size_t i = 0; i < drs.size(); i++) {
                tfo[i] = c.tffield->order(drs[i]);
                if (c.doffield) dofo[i] = c.doffield->order(drs[i]);
            }
            for (auto& g : select_groups(*mesh_, drs, tfo, dofo)) {
                Frag f;
                f.c = &c;
                f.m = m;
                f.type = mesh_->disjreg_type(g[0]);
                f.ptf = c.tffield->order(g[0]);
                f.pdof = c.doffield ? c.doffield->order(g[0]) : f.ptf;
                f.intorder = f.ptf + f.pdof + 2 + c.delta;
                SLX_CHECK(f.intorder >= 0, "trying to integrate at negative order " << f.intorder);
                // hcurl functions are orientation dependent at every order
                const bool hc = c.tffield->hcurl() || (c.doffield && c.doffield->hcurl());
                f.orientdep = hc || h1_orientation_dependent(f.ptf) || (c.doffield && h1_orientation_dependent(f.pdof));
                std::vector<int> sorted = g;
                std::sort(sorted.begin(), sorted.end());
                std::vector<int> elems;
                for (int dr : sorted)
                    for (int e = mesh_->disjreg_begin(dr); e <= mesh_->disjreg_end(dr); e++) elems.push_back(e);
                f.disjregs = sorted;
                if (!f.orientdep) {
                    f.elems = elems;
                    f.orient = mesh_->total_orientation(f.type, elems[0]);
                    frags.push_back(f);
                    continue;
                }
                // elementselector::prepare: stable sort by total orientation, batch by value
                std::vector<int> to(elems.size());
                for (size_t i = 0; i < elems.size(); i++) to[i] = mesh_->total_orientation(f.type, elems[i]);
                std::vector<int> ro(elems.size());
                std::iota(ro.begin(), ro.end(), 0);
                std::stable_sort(ro.begin(), ro.end(), [&](int a, int b) { return to[a] < to[b]; });
                // folded: the batches, in the same order, as one fragment on the
                // orientation-0 tables (element matrices equal up to the signed
                // renumbering, which the addresses and signs carry)
                if (fold_ && !reads_fields(c) && fold_of(basis_of(c.tffield), f.type, f.ptf, f.intorder).ok &&
                    (!c.doffield || fold_of(basis_of(c.doffield), f.type, f.pdof, f.intorder).ok)) {
                    Frag fb = f;
                    fb.orient = 0;
                    fb.folded = true;
                    for (size_t k = 0; k < ro.size(); k++) {
                        fb.elems.push_back(elems[ro[k]]);
                        fb.eorient.push_back(to[ro[k]]);
                    }
                    frags.push_back(fb);
                    continue;
                }
                size_t b = 0;
                while (b < ro.size()) {
                    size_t e = b;
                    while (e < ro.size() && to[ro[e]] == to[ro[b]]) e++;
                    Frag fb = f;
                    fb.orient = to[ro[b]];
                    for (size_t k = b; k < e; k++) fb.elems.push_back(elems[ro[k]]);
                    frags.push_back(fb);
                    b = e;
                }
            }
        }

    // Merge fragments of one integral over the same element batch into buckets.
    std::vector<int> bucket_of(frags.size(), -1);
    std::vector<std::vector<int>> bucket_frags;
    for (size_t i = 0; i < frags.size(); i++) {
        const Frag& f = frags[i];
        for (size_t b = 0; b < bucket_frags.size(); b++) {
            const Frag& g = frags[bucket_frags[b][0]];
            if (g.c->integral_id == f.c->integral_id && g.type == f.type && g.intorder == f.intorder && g.ptf == f.ptf &&
                g.pdof == f.pdof && g.orient == f.orient && g.folded == f.folded && g.elems == f.elems) {
                bucket_of[i] = (int)b;
                break;
            }
        }
        if (bucket_of[i] < 0) {
            bucket_of[i] = (int)bucket_frags.size();
            bucket_frags.emplace_back();
        }
        bucket_frags[bucket_of[i]].push_back((int)i);
    }

    std::vector<std::shared_ptr<RawParameter>> params;
    std::vector<std::shared_ptr<RawField>> fields;
    int nb = (int)bucket_frags.size();
    ps->buckets.resize(nb);
    ps->slices.resize(nb);
    ps->table_ptrs.resize(nb);
    ps->progs.resize(nb);
    int max_elem_id = 0;
    int rank_rhs = 0, rank_mat = 0;
    std::vector<int> frag_rank(frags.size());
    for (size_t i = 0; i < frags.size(); i++) frag_rank[i] = frags[i].m == 0 ? rank_rhs++ : rank_mat++;
    // folded fragment: renumber the [nff][n_elem] address table to the
    // orientation-0 functions and return the signs (same layout)
    auto fold_addresses = [&](const Frag& f, int basis, int order, std::vector<int32_t>& addr) -> const int8_t* {
        const H1Fold& F = fold_of(basis, f.type, order, f.intorder);
        const int ne = (int)f.elems.size(), nff = (int)F.map.size();
        std::vector<int32_t> src = addr;
        ps->signs.emplace_back((size_t)nff * ne, (int8_t)1);
        std::vector<int8_t>& sg = ps->signs.back();
        std::map<int, std::pair<std::vector<int>, std::vector<int8_t>>> cache;
        for (int e = 0; e < ne; e++) {
            auto it = cache.find(f.eorient[e]);
            if (it == cache.end()) {
                std::pair<std::vector<int>, std::vector<int8_t>> ps_;
                h1_fold_apply(F, f.type, f.eorient[e], ps_.first, ps_.second);
                it = cache.emplace(f.eorient[e], std::move(ps_)).first;
            }
            const std::vector<int>& perm = it->second.first;
            const std::vector<int8_t>& sgn = it->second.second;
            for (int i = 0; i < nff; i++) {
                addr[(size_t)perm[i] * ne + e] = src[(size_t)i * ne + e];
                sg[(size_t)perm[i] * ne + e] = sgn[i];
            }
        }
        return sg.data();
    };
    for (int b = 0; b < nb; b++) {
        const Frag& f0 = frags[bucket_frags[b][0]];
        slx_bucket_desc& bd = ps->buckets[b];
        int ne = (int)f0.elems.size();
        bd.elem_type = f0.type;
        bd.n_elem = ne;
        bd.n_geo = mesh_->geometry_nodes(f0.type);  // corners, or every node of a curved element
        ps->ints.emplace_back((size_t)ne * bd.n_geo);
        for (int e = 0; e < ne; e++)
            for (int k = 0; k < bd.n_geo; k++) ps->ints.back()[(size_t)e * bd.n_geo + k] = mesh_->geometry_node(f0.type, f0.elems[e], k);
        bd.geo_nodes = ps->ints.back().data();
        ps->ints.emplace_back(f0.elems.begin(), f0.elems.end());
        bd.elem_ids = ps->ints.back().data();
        for (int e : f0.elems) max_elem_id = std::max(max_elem_id, e + 1);
        GaussRule rule = gauss_rule(f0.type, f0.intorder);
        bd.n_gp = rule.count();
        ps->dbls.push_back(rule.coords);
        bd.gp_coords = ps->dbls.back().data();
        ps->dbls.push_back(rule.weights);
        bd.gp_weights = ps->dbls.back().data();
        // shape tables for the tf and dof orders (orientation of the batch)
        // (family, order): family 0 H1 values and reference derivatives,
        // HCURL_VALUE / HCURL_CURL the hcurl reference components
        std::vector<std::pair<int, int>> table_orders;
        auto table_for = [&](int fam, int order) {
            for (size_t t = 0; t < table_orders.size(); t++)
                if (table_orders[t] == std::make_pair(fam, order)) return (int)t;
            table_orders.push_back({fam, order});
            std::vector<double> tab;
            if (fam == 0) h1_tables(f0.type, order, f0.orient, rule.coords, tab);
            else hcurl_tables(f0.type, order, f0.orient, fam, rule.coords, tab);
            ps->dbls.push_back(tab);
            ps->table_ptrs[b].push_back(ps->dbls.back().data());
            return (int)table_orders.size() - 1;
        };
        Compiler comp;
        comp.consts = &ps->consts;
        comp.params = &params;
        comp.fields = &fields;
        comp.mesh = mesh_;
        comp.disjregs = f0.disjregs;
        int ncoefs = 0;
        for (int fi : bucket_frags[b]) {
            const Frag& f = frags[fi];
            const Contribution& c = *f.c;
            slx_slice_desc sd{};
            sd.frag_rank = frag_rank[fi];
            sd.is_rhs = c.doffield ? 0 : 1;
            sd.tf_table = table_for(c.tffam, f.ptf);
            sd.n_tf = element_nff(*c.tffield, f.type, f.ptf);
            ps->ints.push_back(dofm_->getaddresses(c.tffield, f.ptf, f.type, f.elems, c.tfphysreg));
            sd.tf_addr = ps->ints.back().data();
            sd.tf_sign = nullptr;
            sd.dof_sign = nullptr;
            if (f.folded && (c.tffield->hcurl() || h1_orientation_dependent(f.ptf)))
                sd.tf_sign = fold_addresses(f, basis_of(c.tffield), f.ptf, ps->ints.back());
            if (c.doffield) {
                sd.dof_table = table_for(c.doffam, f.pdof);
                sd.n_dof = element_nff(*c.doffield, f.type, f.pdof);
                ps->ints.push_back(dofm_->getaddresses(c.doffield, f.pdof, f.type, f.elems, c.dofphysreg));
                sd.dof_addr = ps->ints.back().data();
                if (f.folded && (c.doffield->hcurl() || h1_orientation_dependent(f.pdof)))
                    sd.dof_sign = fold_addresses(f, basis_of(c.doffield), f.pdof, ps->ints.back());
            } else {
                sd.dof_table = -1;
                sd.n_dof = 1;
                sd.dof_addr = nullptr;
            }
            ps->terms.emplace_back();
            for (size_t t = 0; t < c.tfs.size(); t++) {
                OpP co = op_simplify(c.coeffs[t]->copy());
                int v = comp.compile(co);
                slx_instr out;
                out.op = SLX_OP_OUT;
                out.dst = 0;
                out.a = (int16_t)v;
                out.b = 0;
                out.imm = ncoefs;
                comp.code.push_back(out);
                slx_term term;
                term.tf_deriv = c.tfs[t]->kietaphideriv;
                term.dof_deriv = c.dofs[t] ? c.dofs[t]->kietaphideriv : 0;
                term.coef = ncoefs++;
                ps->terms.back().push_back(term);
            }
            sd.n_terms = (int)ps->terms.back().size();
            sd.terms = ps->terms.back().data();
            ps->slices[b].push_back(sd);
        }
        bd.n_slices = (int)ps->slices[b].size();
        bd.slices = ps->slices[b].data();
        ps->ints.emplace_back();
        for (auto& fo : table_orders) ps->ints.back().push_back(fo.first ? hcurl_count(f0.type, fo.second) : h1_count(f0.type, fo.second));
        bd.n_tables = (int)table_orders.size();
        bd.table_nff = ps->ints.back().data();
        bd.tables = ps->table_ptrs[b].data();
        ps->progs[b] = comp.code;
        bd.n_regs = allocate_registers(ps->progs[b]);
        bd.prog = ps->progs[b].data();
        bd.n_instr = (int)ps->progs[b].size();
        bd.n_coefs = ncoefs;
        bool affine = f0.type == TRI || f0.type == TET || f0.type == LINE;
        bd.q_invariant = (!comp.uses_q && affine) ? 1 : 0;
    }
    // numeric parameter arrays
    ps->n_params = (int)params.size();
    ps->param_len = max_elem_id;
    ps->params.assign((size_t)ps->n_params * max_elem_id, 0.0);
    for (int k = 0; k < ps->n_params; k++) {
        const std::vector<double>& v = params[k]->element_values;
        SLX_CHECK((int64_t)v.size() >= max_elem_id, "per-element parameter has " << v.size() << " values, needs " << max_elem_id);
        std::copy(v.begin(), v.begin() + max_elem_id, ps->params.begin() + (size_t)k * max_elem_id);
    }
    // fields read by coefficient programs (Newton re-assembly)
    ps->plan_fields = fields;
    int nf = (int)fields.size();
    ps->field_nff.assign((size_t)nf * nb, 0);
    ps->field_addr.assign((size_t)nf * nb, nullptr);
    ps->field_tables.assign((size_t)nf * nb, nullptr);
    for (int k = 0; k < nf; k++) {
        auto& fl = fields[k];
        if (fl->data_size() == 0) fl->build_data_numbering();
        for (int b = 0; b < nb; b++) {
            const Frag& f0 = frags[bucket_frags[b][0]];
            const slx_bucket_desc& bd = ps->buckets[b];
            int order = fl->order(f0.disjregs[0]);
            std::vector<FormFunction> ffl = h1_iterate(f0.type, order);
            std::vector<int32_t> addr((size_t)ffl.size() * bd.n_elem);
            for (size_t q = 0; q < ffl.size(); q++)
                for (int i = 0; i < bd.n_elem; i++) {
                    int e = f0.elems[i];
                    int sub = mesh_->subelement(ffl[q].assoc_type, f0.type, e, ffl[q].sub);
                    int dr = mesh_->disjreg_of(ffl[q].assoc_type, sub);
                    int cnt = h1_count(ffl[q].assoc_type, fl->order(dr), ffl[q].dim);
                    addr[q * bd.n_elem + i] = ffl[q].index < cnt ? fl->data_rangebegin[dr][ffl[q].index] + (sub - mesh_->disjreg_begin(dr)) : -1;
                }
            ps->ints.push_back(addr);
            ps->field_addr[(size_t)k * nb + b] = ps->ints.back().data();
            ps->field_nff[(size_t)k * nb + b] = (int)ffl.size();
            std::vector<double> tab;
            GaussRule rule = gauss_rule(f0.type, f0.intorder);
            h1_tables(f0.type, order, f0.orient, rule.coords, tab);
            ps->dbls.push_back(tab);
            ps->field_tables[(size_t)k * nb + b] = ps->dbls.back().data();
        }
        ps->field_data_ptrs.push_back(fl->data.data());
        ps->field_len.push_back(fl->data_size());
    }
    ps->constrained = dofm_->isconstrained();
    ps->dirichlet = dofm_->constraint_values();
    slx_plan_desc& d = ps->desc;
    d.mesh_dim = mesh_->dimension();
    d.n_nodes = mesh_->count(0);
    d.n_dofs = dofm_->countdofs();
    d.constrained = ps->constrained.data();
    d.n_consts = (int)ps->consts.size();
    d.consts = ps->consts.data();
    d.n_buckets = nb;
    d.buckets = ps->buckets.data();
    d.n_fields = nf;
    d.field_nff = ps->field_nff.data();
    d.field_addr = ps->field_addr.data();
    d.field_tables = ps->field_tables.data();
    d.has_rhs = !contribs_[0].empty();
    d.has_matrix = !contribs_[1].empty();
    d.dirichlet = ps->dirichlet.data();
    for (auto& b : ps->buckets) ps->n_elem_total += b.n_elem;
    return ps;
}

}  // namespace slx
