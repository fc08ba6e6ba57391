// Symbolic integrand expressions and the dof/tf term extraction.
//
// The operation tree mirrors the reference's operation classes
// (src/expression/operation/op*.cpp) for the node kinds the assembly path
// needs: constants, parameters, dof/tf/field values and space derivatives,
// Jacobian entries, sums, products, inversion, power and a few unary
// functions.  expand / group / simplify and extractdoftf follow
// opsum.cpp:80-145, opproduct.cpp:121-240 and expression.cpp:2105-2296 so that
// slices, entries and their order match the reference.
#pragma once
#include <memory>
#include <string>
#include <vector>
#include "common.h"

namespace slx {

class RawField;
class RawParameter;

enum class OpKind {
    Constant, Parameter, Dof, Tf, Field, InvJac, DetJac, Jac, Coord,
    Sum, Product, Inversion, Power, Abs, Exp, Log, Sin, Cos, Tan, Sqrt
};

struct Op;
using OpP = std::shared_ptr<Op>;

struct Op : std::enable_shared_from_this<Op> {
    OpKind kind = OpKind::Constant;
    double value = 0.0;                  // Constant
    std::shared_ptr<RawField> field;     // Dof / Tf / Field (a scalar component field)
    std::shared_ptr<RawParameter> param; // Parameter (scalar)
    int spacederiv = 0;                  // 0 none, 1 d/dx, 2 d/dy, 3 d/dz
    int kietaphideriv = 0;               // 0 none, 1 d/dki, 2 d/deta, 3 d/dphi
                                         // (hcurl ops: the reference component 1..3)
    int hc = 0;                          // Dof / Tf / Field of an hcurl field: 1 the
                                         // function value, 2 its reference curl
    int physreg = -1;                    // dof(u, physreg) / tf(u, physreg)
    int row = 0, col = 0;                // Jac / InvJac entry, Coord axis
    bool reuse = false;
    std::vector<OpP> args;

    bool isconstant() const { return kind == OpKind::Constant; }
    bool issum() const { return kind == OpKind::Sum; }
    bool isproduct() const { return kind == OpKind::Product; }
    bool isdof() const { return kind == OpKind::Dof; }
    bool istf() const { return kind == OpKind::Tf; }
    bool isdofincluded() const;
    bool istfincluded() const;
    OpP copy() const;   // shallow copy like operation::copy (arguments shared)
};

OpP make_const(double v);
OpP make_op(OpKind k, std::vector<OpP> args);

// Tree transformations with the reference semantics.
OpP op_expand(const OpP& op);
void op_group(const OpP& op);
OpP op_simplify(const OpP& op);
std::string op_str(const OpP& op);

// Result of expression::extractdoftf: per slice, per entry.
struct DofTfSlices {
    std::vector<std::vector<OpP>> coeffs, dofs, tfs;
};
DofTfSlices extract_dof_tf(OpP expr, int elementdimension);

// A (rows x cols) array of operations (expression.h).
class expression {
   public:
    expression() : rows_(1), cols_(1), ops_{make_const(0.0)} {}
    expression(double v) : rows_(1), cols_(1), ops_{make_const(v)} {}
    explicit expression(OpP op) : rows_(1), cols_(1), ops_{op} {}
    expression(int rows, int cols, const std::vector<expression>& entries);

    int countrows() const { return rows_; }
    int countcolumns() const { return cols_; }
    bool isscalar() const { return rows_ == 1 && cols_ == 1; }
    expression at(int r, int c) const;
    const OpP& op(int r = 0, int c = 0) const { return ops_[r * cols_ + c]; }
    OpP& op(int r = 0, int c = 0) { return ops_[r * cols_ + c]; }
    expression getcopy() const;
    expression transpose() const;
    void reuseit(bool flag = true);

    expression operator-() const;
    friend expression operator+(const expression& a, const expression& b);
    friend expression operator-(const expression& a, const expression& b);
    friend expression operator*(const expression& a, const expression& b);
    friend expression operator/(const expression& a, const expression& b);

    // space derivative 1..3 applied to every entry (expression::spacederivative)
    expression spacederivative(int whichderivative) const;
    expression dof(int physreg) const;
    expression tf(int physreg) const;

    // An hcurl field, or its dof() / tf(), keeps the field it maps from the
    // reference element (expression::inrefcoord, expression.cpp:30-35): curl()
    // is then taken in reference coordinates (sl.cpp:1184-1222).  Every other
    // operation returns a plain expression.
    const std::shared_ptr<RawField>& hcurl_field() const { return hc_field_; }
    OpKind hcurl_kind() const { return hc_kind_; }
    int hcurl_physreg() const { return hc_physreg_; }
    void set_hcurl(const std::shared_ptr<RawField>& f, OpKind k, int physreg) {
        hc_field_ = f;
        hc_kind_ = k;
        hc_physreg_ = physreg;
    }

   private:
    int rows_, cols_;
    std::vector<OpP> ops_;
    std::shared_ptr<RawField> hc_field_;
    OpKind hc_kind_ = OpKind::Field;
    int hc_physreg_ = -1;
};

}  // namespace slx
