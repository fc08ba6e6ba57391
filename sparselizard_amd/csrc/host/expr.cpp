// Expression trees, expansion and dof/tf term extraction (see expr.h).
#include "expr.h"
#include <cmath>
#include "fem.h"

namespace slx {

OpP make_const(double v) {
    auto o = std::make_shared<Op>();
    o->kind = OpKind::Constant;
    o->value = v;
    return o;
}

OpP make_op(OpKind k, std::vector<OpP> args) {
    auto o = std::make_shared<Op>();
    o->kind = k;
    o->args = std::move(args);
    return o;
}

bool Op::isdofincluded() const {
    if (kind == OpKind::Dof) return true;
    for (auto& a : args)
        if (a->isdofincluded()) return true;
    return false;
}

bool Op::istfincluded() const {
    if (kind == OpKind::Tf) return true;
    for (auto& a : args)
        if (a->istfincluded()) return true;
    return false;
}

OpP Op::copy() const {
    auto o = std::make_shared<Op>(*this);
    o->reuse = false;
    return o;
}

// ---- group / expand / simplify (opsum.cpp:80-145, opproduct.cpp:121-240) ----

void op_group(const OpP& op) {
    if (op->kind != OpKind::Sum && op->kind != OpKind::Product) {
        for (auto& a : op->args) op_group(a);
        return;
    }
    std::vector<OpP> g;
    for (auto& a : op->args) {
        op_group(a);
        if (a->kind == op->kind && !a->reuse)
            g.insert(g.end(), a->args.begin(), a->args.end());
        else
            g.push_back(a);
    }
    op->args = g;
}

OpP op_expand(const OpP& op) {
    if (op->kind == OpKind::Sum) {
        for (auto& a : op->args)
            if (a->isdofincluded() || a->istfincluded()) a = op_expand(a);
        op_group(op);
        return op;
    }
    if (op->kind != OpKind::Product) return op;
    for (auto& a : op->args)
        if (a->isdofincluded() || a->istfincluded()) a = op_expand(a);
    op_group(op);
    // Every factor as a sum (possibly of a single term); sums without dof/tf
    // stay monolithic (opproduct.cpp:137-147).
    std::vector<OpP> prod(op->args.size());
    for (size_t i = 0; i < op->args.size(); i++) {
        const OpP& a = op->args[i];
        if (!a->issum() || (!a->isdofincluded() && !a->istfincluded()))
            prod[i] = make_op(OpKind::Sum, {a});
        else
            prod[i] = a;
    }
    OpP expanded = prod[0];
    for (size_t i = 1; i < prod.size(); i++) {
        OpP s = make_op(OpKind::Sum, {});
        for (auto& ej : expanded->args)
            for (auto& pk : prod[i]->args) {
                OpP p = make_op(OpKind::Product, {ej, pk});
                op_group(p);
                s->args.push_back(p);
            }
        expanded = s;
    }
    if (expanded->args.size() <= 1) expanded = expanded->args[0];
    return expanded;
}

OpP op_simplify(const OpP& op) {
    if (op->kind == OpKind::Sum) {
        for (auto& a : op->args) a = op_simplify(a);
        op_group(op);
        double summed = 0;
        for (int i = (int)op->args.size() - 1; i >= 0; i--)
            if (op->args[i]->isconstant()) {
                summed += op->args[i]->value;
                op->args.erase(op->args.begin() + i);
            }
        if (!op->args.empty()) {
            if (summed != 0) op->args.push_back(make_const(summed));
            return op;
        }
        op->args = {make_const(summed)};
        return op->args[0];
    }
    if (op->kind == OpKind::Product) {
        for (auto& a : op->args) a = op_simplify(a);
        op_group(op);
        double mult = 1;
        for (int i = (int)op->args.size() - 1; i >= 0; i--)
            if (op->args[i]->isconstant()) {
                double v = op->args[i]->value;
                if (v == 0) {
                    op->args = {make_const(0)};
                    return op->args[0];
                }
                mult *= v;
                op->args.erase(op->args.begin() + i);
            }
        if (!op->args.empty()) {
            if (mult != 1) op->args.push_back(make_const(mult));
            return op;
        }
        op->args = {make_const(mult)};
        return op->args[0];
    }
    for (auto& a : op->args) a = op_simplify(a);
    // Constant folding of unary/binary functions of constants (op*::simplify).
    bool allconst = !op->args.empty();
    for (auto& a : op->args) allconst = allconst && a->isconstant();
    if (allconst) {
        double x = op->args[0]->value;
        switch (op->kind) {
            case OpKind::Inversion: return make_const(1.0 / x);
            case OpKind::Power: return make_const(std::pow(x, op->args[1]->value));
            case OpKind::Abs: return make_const(std::abs(x));
            case OpKind::Exp: return make_const(std::exp(x));
            case OpKind::Log: return make_const(std::log(x));
            case OpKind::Sin: return make_const(std::sin(x));
            case OpKind::Cos: return make_const(std::cos(x));
            case OpKind::Tan: return make_const(std::tan(x));
            case OpKind::Sqrt: return make_const(std::sqrt(x));
            default: break;
        }
    }
    return op;
}

std::string op_str(const OpP& op) {
    std::ostringstream os;
    switch (op->kind) {
        case OpKind::Constant: os << op->value; break;
        case OpKind::Parameter: os << "param"; break;
        case OpKind::Dof: case OpKind::Tf: case OpKind::Field: {
            const char* nm = op->kind == OpKind::Dof ? "dof" : op->kind == OpKind::Tf ? "tf" : "field";
            os << nm << "(" << (op->field ? op->field->name() : "?") << ")";
            if (op->spacederiv) os << "_d" << "xyz"[op->spacederiv - 1];
            if (op->kietaphideriv) os << "_d" << "kep"[op->kietaphideriv - 1];
            break;
        }
        case OpKind::InvJac: os << "invjac(" << op->row << "," << op->col << ")"; break;
        case OpKind::Jac: os << "jac(" << op->row << "," << op->col << ")"; break;
        case OpKind::DetJac: os << "detjac"; break;
        case OpKind::Coord: os << "xyz"[op->row]; break;
        case OpKind::Sum: case OpKind::Product: {
            os << "(";
            for (size_t i = 0; i < op->args.size(); i++) os << (i ? (op->kind == OpKind::Sum ? " + " : " * ") : "") << op_str(op->args[i]);
            os << ")";
            break;
        }
        case OpKind::Inversion: os << "1/" << op_str(op->args[0]); break;
        case OpKind::Power: os << "pow(" << op_str(op->args[0]) << "," << op_str(op->args[1]) << ")"; break;
        default: os << "f(" << op_str(op->args[0]) << ")"; break;
    }
    return os.str();
}

// ---- extractdoftf (expression.cpp:2105-2296) ----

DofTfSlices extract_dof_tf(OpP root, int elementdimension) {
    root = op_simplify(root);
    DofTfSlices out;
    if (root->isconstant() && root->value == 0) return out;
    std::vector<OpP> sumterms = root->issum() ? root->args : std::vector<OpP>{root};
    for (OpP term : sumterms) {
        if (term->istf()) term = make_op(OpKind::Product, {term});
        SLX_CHECK(term->isproduct(), "don't know what to do with the expression provided to the formulation; "
                                     "it should be rewritable into a sum of products coef*dof*tf");
        OpP curdof, curtf;
        OpP coef = term->copy();
        for (int j = (int)term->args.size() - 1; j >= 0; j--) {
            const OpP& a = term->args[j];
            if (!curdof && a->isdof()) {
                curdof = a;
                coef->args.erase(coef->args.begin() + j);
            }
            if (!curtf && a->istf()) {
                curtf = a;
                coef->args.erase(coef->args.begin() + j);
            }
        }
        if (coef->args.empty()) coef = make_const(1);
        SLX_CHECK(!coef->isdofincluded() && !coef->istfincluded(),
                  "don't know what to do with the expression provided to the formulation (dof or tf inside the coefficient)");
        SLX_CHECK(curtf, "malformed expression provided to the formulation (the test function is missing)");
        // slice: same dof field, tf field, physical regions (time derivatives not modelled)
        int slice = -1;
        for (size_t s = 0; s < out.tfs.size(); s++) {
            const OpP& d0 = out.dofs[s][0];
            const OpP& t0 = out.tfs[s][0];
            auto dfield = [](const OpP& o) { return o ? o->field.get() : nullptr; };
            auto dphys = [](const OpP& o) { return o ? o->physreg : -1; };
            auto dfam = [](const OpP& o) { return o ? o->hc : 0; };
            // hcurl values and curls use different tables: separate slices
            if (dfield(d0) == dfield(curdof) && t0->field.get() == curtf->field.get() && dphys(d0) == dphys(curdof) &&
                t0->physreg == curtf->physreg && dfam(d0) == dfam(curdof) && t0->hc == curtf->hc) {
                slice = (int)s;
                break;
            }
        }
        if (slice < 0) {
            slice = (int)out.tfs.size();
            out.coeffs.emplace_back();
            out.dofs.emplace_back();
            out.tfs.emplace_back();
        }
        int dsd = curdof ? curdof->spacederiv : 0;
        int numdof = dsd > 0 ? std::max(1, elementdimension) : 1;
        int numtf = curtf->spacederiv > 0 ? std::max(1, elementdimension) : 1;
        std::vector<OpP> dsplit(numdof * numtf), tsplit(numdof * numtf), csplit(numdof * numtf);
        for (int i = 0; i < numdof; i++)
            for (int j = 0; j < numtf; j++) {
                OpP cd = curdof ? curdof->copy() : nullptr;
                OpP ct = curtf->copy();
                OpP nc = coef;
                coef->reuse = true;
                if (dsd > 0) {
                    auto ij = std::make_shared<Op>();
                    ij->kind = OpKind::InvJac;
                    ij->row = dsd - 1;
                    ij->col = i;
                    nc = make_op(OpKind::Product, {nc, ij});
                    cd->kietaphideriv = i + 1;
                    cd->spacederiv = 0;
                }
                if (curtf->spacederiv > 0) {
                    auto ij = std::make_shared<Op>();
                    ij->kind = OpKind::InvJac;
                    ij->row = curtf->spacederiv - 1;
                    ij->col = j;
                    nc = make_op(OpKind::Product, {nc, ij});
                    ct->kietaphideriv = j + 1;
                    ct->spacederiv = 0;
                }
                dsplit[i * numtf + j] = cd;
                tsplit[i * numtf + j] = ct;
                bool hasderiv = dsd > 0 || curtf->spacederiv > 0;
                csplit[i * numtf + j] = (elementdimension > 0 || !hasderiv) ? nc : make_const(0);
            }
        for (size_t t = 0; t < tsplit.size(); t++) {
            int entry = -1;
            for (size_t e = 0; e < out.tfs[slice].size(); e++) {
                const OpP& de = out.dofs[slice][e];
                const OpP& te = out.tfs[slice][e];
                int dk = de ? de->kietaphideriv : 0, ndk = dsplit[t] ? dsplit[t]->kietaphideriv : 0;
                if (dk == ndk && te->kietaphideriv == tsplit[t]->kietaphideriv) {
                    entry = (int)e;
                    out.coeffs[slice][e] = make_op(OpKind::Sum, {out.coeffs[slice][e], csplit[t]});
                    break;
                }
            }
            if (entry < 0) {
                out.coeffs[slice].push_back(csplit[t]);
                out.dofs[slice].push_back(dsplit[t]);
                out.tfs[slice].push_back(tsplit[t]);
            }
        }
    }
    return out;
}

// ---- expression arrays (expression.cpp) ----

expression::expression(int rows, int cols, const std::vector<expression>& entries) : rows_(rows), cols_(cols) {
    SLX_CHECK((int)entries.size() == rows * cols || ((int)entries.size() == rows && rows == cols),
              "expression: expected " << rows * cols << " entries");
    ops_.resize(rows * cols);
    if ((int)entries.size() == rows * cols) {
        for (int i = 0; i < rows * cols; i++) {
            SLX_CHECK(entries[i].isscalar(), "expression entries must be scalar");
            ops_[i] = entries[i].op();
        }
    } else {
        // diagonal form (expression.cpp:61)
        for (int i = 0; i < rows * cols; i++) ops_[i] = make_const(0);
        for (int i = 0; i < rows; i++) ops_[i * cols + i] = entries[i].op();
    }
}

expression expression::at(int r, int c) const {
    SLX_CHECK(r >= 0 && r < rows_ && c >= 0 && c < cols_, "trying to get entry (" << r << ", " << c << ") in a " << rows_ << " by " << cols_ << " expression array");
    return expression(ops_[r * cols_ + c]);
}

expression expression::getcopy() const { return *this; }

expression expression::transpose() const {
    expression o = *this;
    o.hc_field_.reset();
    o.rows_ = cols_;
    o.cols_ = rows_;
    for (int r = 0; r < rows_; r++)
        for (int c = 0; c < cols_; c++) o.ops_[c * rows_ + r] = ops_[r * cols_ + c];
    return o;
}

void expression::reuseit(bool flag) {
    for (auto& o : ops_) o->reuse = flag;
}

expression operator+(const expression& a, const expression& b) {
    SLX_CHECK(a.rows_ == b.rows_ && a.cols_ == b.cols_, "trying to add a " << a.rows_ << "x" << a.cols_ << " expression to a " << b.rows_ << "x" << b.cols_);
    expression o = a;
    o.hc_field_.reset();
    for (size_t i = 0; i < o.ops_.size(); i++) o.ops_[i] = make_op(OpKind::Sum, {a.ops_[i], b.ops_[i]});
    return o;
}

expression operator*(const expression& a, const expression& b) {
    // expression.cpp:2399-2459
    if (a.isscalar()) {
        expression o = b;
        o.hc_field_.reset();
        for (size_t i = 0; i < o.ops_.size(); i++) o.ops_[i] = make_op(OpKind::Product, {a.ops_[0], b.ops_[i]});
        return o;
    }
    if (b.isscalar()) {
        expression o = a;
        o.hc_field_.reset();
        for (size_t i = 0; i < o.ops_.size(); i++) o.ops_[i] = make_op(OpKind::Product, {a.ops_[i], b.ops_[0]});
        return o;
    }
    if ((a.rows_ == 1 || a.cols_ == 1) && a.rows_ == b.rows_ && a.cols_ == b.cols_) {
        OpP s = make_op(OpKind::Sum, {});
        for (int i = 0; i < std::max(a.rows_, a.cols_); i++) s->args.push_back(make_op(OpKind::Product, {a.ops_[i], b.ops_[i]}));
        return expression(s);
    }
    SLX_CHECK(a.cols_ == b.rows_, "trying to multiply a " << a.rows_ << "x" << a.cols_ << " expression by a " << b.rows_ << "x" << b.cols_);
    expression o;
    o.rows_ = a.rows_;
    o.cols_ = b.cols_;
    o.ops_.resize(o.rows_ * o.cols_);
    for (int r = 0; r < a.rows_; r++)
        for (int c = 0; c < b.cols_; c++) {
            OpP s = make_op(OpKind::Sum, {});
            for (int k = 0; k < a.cols_; k++) s->args.push_back(make_op(OpKind::Product, {a.ops_[r * a.cols_ + k], b.ops_[k * b.cols_ + c]}));
            o.ops_[r * o.cols_ + c] = s;
        }
    return o;
}

expression expression::operator-() const { return (*this) * expression(-1.0); }
expression operator-(const expression& a, const expression& b) { return a + b * expression(-1.0); }

expression operator/(const expression& a, const expression& b) {
    SLX_CHECK(b.isscalar(), "only divisions by scalars are allowed");
    expression o = a;
    o.hc_field_.reset();
    for (size_t i = 0; i < o.ops_.size(); i++) o.ops_[i] = make_op(OpKind::Product, {a.ops_[i], make_op(OpKind::Inversion, {b.ops_[0]})});
    return o;
}

expression expression::spacederivative(int which) const {
    SLX_CHECK(rows_ <= 3 && cols_ == 1, "can only take space derivatives of column vectors with up to 3 components");
    SLX_CHECK(!hc_field_, "space derivatives of hcurl fields are not supported (use curl)");
    expression o = *this;
    for (auto& op : o.ops_) {
        if (op->isconstant()) {
            op = make_const(0);
            continue;
        }
        SLX_CHECK(op->kind == OpKind::Field || op->kind == OpKind::Dof || op->kind == OpKind::Tf,
                  "space derivatives can only be applied to fields, dof() and tf()");
        SLX_CHECK(op->spacederiv == 0, "only first order space derivatives are supported");
        if (which > op->field->meshdim()) {
            op = make_op(OpKind::Product, {op, make_const(0)});
        } else {
            op = op->copy();
            op->spacederiv = which;
        }
    }
    return o;
}

// Copy of an op tree with every hcurl leaf turned into a Dof / Tf leaf.
static OpP retag_hcurl(const OpP& op, OpKind k, int physreg) {
    if (op->kind == OpKind::Field && op->hc) {
        OpP n = op->copy();
        n->kind = k;
        n->physreg = physreg;
        return n;
    }
    if (op->args.empty()) return op;
    OpP n = op->copy();
    for (auto& a : n->args) a = retag_hcurl(a, k, physreg);
    return n;
}

static expression tag(const expression& in, OpKind k, int physreg) {
    if (in.hcurl_field()) {
        SLX_CHECK(in.hcurl_kind() == OpKind::Field, "dof()/tf() of a dof()/tf() expression");
        expression o = in;
        for (int r = 0; r < in.countrows(); r++)
            for (int c = 0; c < in.countcolumns(); c++) o.op(r, c) = retag_hcurl(in.op(r, c), k, physreg);
        o.set_hcurl(in.hcurl_field(), k, physreg);
        return o;
    }
    expression o = in;
    for (int r = 0; r < in.countrows(); r++)
        for (int c = 0; c < in.countcolumns(); c++) {
            OpP& op = o.op(r, c);
            if (op->isconstant() && op->value == 0) continue;
            SLX_CHECK(op->kind == OpKind::Field, "the argument of dof()/tf() must be a field or a field expression (constant 0 is allowed)");
            SLX_CHECK(op->spacederiv == 0 && op->kietaphideriv == 0, "cannot apply space or time derivatives to the dof()/tf() field argument");
            OpP n = op->copy();
            n->kind = k;
            n->physreg = physreg;
            op = n;
        }
    return o;
}

expression expression::dof(int physreg) const { return tag(*this, OpKind::Dof, physreg); }
expression expression::tf(int physreg) const { return tag(*this, OpKind::Tf, physreg); }

}  // namespace slx
