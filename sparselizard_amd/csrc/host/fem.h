// Host-side objects mirroring the reference's user API for the assembly
// path: fields, parameters, the DOF manager and the formulation with
// generate() / A() / b().  generate() builds an assembly plan description
// (slx_plan_desc, include/slx.h) and hands the numeric work to the GPU.
//
// Reference anchors:
//   field / rawfield       src/field/field.cpp, src/field/rawfield.cpp:652-722
//   dofmanager             src/formulation/dofmanager.cpp:35-123, 366-451, 846-918
//   formulation            src/formulation/formulation.cpp:51-149, 176-218, 358-422
//   contribution::generate src/formulation/contribution.cpp:20-307
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>
#include "../../../include/slx.h"
#include "expr.h"
#include "mesh.h"

namespace slx {

class RawField {
   public:
    RawField(const Mesh* m, const std::string& type, const std::string& name);
    const std::string& name() const { return name_; }
    const std::string& type() const { return type_; }
    bool hcurl() const { return type_ == "hcurl"; }
    int meshdim() const { return mesh_->dimension(); }
    const Mesh* mesh() const { return mesh_; }
    int ncomponents() const { return subfields_.empty() ? 1 : (int)subfields_.size(); }
    std::shared_ptr<RawField> component(int i) { return subfields_.empty() ? self() : subfields_[i]; }
    int order(int disjreg) const { return orders_[disjreg]; }
    const std::vector<int>& orders() const { return orders_; }
    bool constrained(int disjreg) const { return constrained_[disjreg] != 0; }
    double constraint_value(int disjreg) const { return cvalue_[disjreg]; }
    // rawfield::setorder (rawfield.cpp:652-722): all disjoint regions of the
    // physical region (every dimension) get the order; later calls override.
    void setorder(int physreg, int order);
    // rawfield::setdisjregconstraint with a constant value (rawfield.cpp:1166-1248).
    void setconstraint(int physreg, double value = 0.0);
    // rawfield::setdisjregconstraint with an expression (hcurl fields: a 3x1
    // vector, projected on the constrained edges when the DOF values are
    // computed; later calls override earlier ones region by region).
    void setconstraint(int physreg, const expression& value, int extraintegrationdegree = 0);
    const expression* constraint_expression(int disjreg) const { return cexpr_[disjreg].get(); }
    int constraint_integration_degree(int disjreg) const { return ceio_[disjreg]; }
    std::shared_ptr<RawField> self() { return self_.lock(); }
    void setself(const std::shared_ptr<RawField>& s) { self_ = s; }
    void addsub(const std::shared_ptr<RawField>& s) { subfields_.push_back(s); }
    // Field data (coefficients in the field's own numbering: all disjoint
    // regions, in order, like a dofmanager holding only this field).
    std::vector<double> data;
    std::vector<std::vector<int>> data_rangebegin;  // [disjreg][ff]
    int64_t data_size() const { return data_size_; }
    void build_data_numbering();

   private:
    const Mesh* mesh_;
    std::string type_, name_;
    std::vector<std::shared_ptr<RawField>> subfields_;
    std::vector<int> orders_;
    std::vector<char> constrained_;
    std::vector<double> cvalue_;
    std::vector<std::shared_ptr<expression>> cexpr_;
    std::vector<int> ceio_;
    std::weak_ptr<RawField> self_;
    int64_t data_size_ = 0;
};

// User handle: field u(mesh, "h1") / field u(mesh, "h1xyz").
class field {
   public:
    field() = default;
    field(const Mesh& m, const std::string& type, const std::string& name = "u");
    void setorder(int physreg, int order);
    void setconstraint(int physreg, double value = 0.0);
    void setconstraint(int physreg, const expression& value, int extraintegrationdegree = 0);
    std::shared_ptr<RawField> raw() const { return raw_; }
    operator expression() const;
    expression comp(int i) const;

   private:
    std::shared_ptr<RawField> raw_;
};

// Scalar parameter: a constant per physical region, or per-element values
// (an array indexed by element number for the integration element type).
class RawParameter {
   public:
    std::map<int, double> region_values;   // physreg -> value
    std::vector<double> element_values;    // optional per-element values
    int index = -1;                        // parameter slot in the numeric inputs
};

class parameter {
   public:
    parameter() : raw_(std::make_shared<RawParameter>()) {}
    void set(int physreg, double v) { raw_->region_values[physreg] = v; }
    void setelementvalues(const std::vector<double>& v) { raw_->element_values = v; }
    std::shared_ptr<RawParameter> raw() const { return raw_; }
    operator expression() const;

   private:
    std::shared_ptr<RawParameter> raw_;
};

struct integration {
    int physreg;
    expression expr;
    int delta;
};

// Free functions of the sl namespace used to write integrands.
namespace sl {
expression dof(const expression& e, int physreg = -1);
expression tf(const expression& e, int physreg = -1);
expression dx(const expression& e);
expression dy(const expression& e);
expression dz(const expression& e);
expression grad(const expression& e);
// sl::curl (sl.cpp:1167-1226): of an hcurl field (or its dof/tf) in reference
// coordinates, mapped by J^T / det J; else from the space derivatives.
expression curl(const expression& e);
expression strain(const expression& e);
expression predefinedelasticity(const expression& dofu, const expression& tfu, const expression& E, const expression& nu);
expression pow(const expression& b, const expression& e);
expression abs(const expression& e);
expression sqrt(const expression& e);
expression exp(const expression& e);
expression sin(const expression& e);
expression cos(const expression& e);
expression norm(const expression& e);
expression x();
expression y();
expression z();
integration integral(int physreg, const expression& e, int delta = 0);
}  // namespace sl

// DOF numbering (dofmanager.cpp).
class DofManager {
   public:
    explicit DofManager(const Mesh* m) : mesh_(m) {}
    void addtostructure(const std::shared_ptr<RawField>& f, int physreg);
    void addtostructure(const std::shared_ptr<RawField>& f, const std::vector<int>& disjregs);
    int64_t countdofs() const { return ndofs_; }
    std::vector<uint8_t> isconstrained() const;
    // Dirichlet values of the constrained DOFs (0 elsewhere).
    std::vector<double> constraint_values() const;
    // sl::setdata(vec): copies the values of a solution vector (this DOF
    // numbering) into the data of every field of the structure that holds
    // data (the field's own numbering).
    void setdata(const double* x, int64_t n) const;
    // dofmanager::getaddresses: [n_ff][elems.size()] row-major.
    std::vector<int32_t> getaddresses(const std::shared_ptr<RawField>& f, int order, int elemtype,
                                      const std::vector<int>& elems, int fieldphysreg) const;
    int fieldindex(const RawField* f) const;
    const std::vector<std::vector<int>>& rangebegin(int fi) const { return rangebegin_[fi]; }

   private:
    const Mesh* mesh_;
    std::vector<std::shared_ptr<RawField>> fields_;
    std::vector<std::vector<int>> orders_at_add_;
    std::vector<std::vector<std::vector<int>>> rangebegin_;  // [field][disjreg][ff]
    int64_t ndofs_ = 0;
};

// A plan description plus the storage its pointers refer to.
struct PlanStorage {
    slx_plan_desc desc{};
    std::vector<slx_bucket_desc> buckets;
    std::vector<std::vector<slx_slice_desc>> slices;
    std::vector<std::vector<slx_term>> terms;                 // per (bucket, slice) flattened
    std::vector<std::vector<int32_t>> ints;                   // owned int arrays
    std::vector<std::vector<int8_t>> signs;                   // owned sign arrays
    std::vector<std::vector<double>> dbls;                    // owned double arrays
    std::vector<std::vector<const double*>> table_ptrs;
    std::vector<std::vector<slx_instr>> progs;
    std::vector<double> consts;
    std::vector<uint8_t> constrained;
    std::vector<double> dirichlet;
    std::vector<int32_t> field_nff;
    std::vector<const int32_t*> field_addr;
    std::vector<const double*> field_tables;
    // numeric inputs
    std::vector<double> params;                               // [n_params][param_len]
    int64_t param_len = 0;
    int32_t n_params = 0;
    std::vector<std::shared_ptr<RawField>> plan_fields;       // fields read by SLX_OP_FIELD
    std::vector<const double*> field_data_ptrs;
    std::vector<int64_t> field_len;
    int64_t n_elem_total = 0;
};

class formulation {
   public:
    explicit formulation(const Mesh& m);
    formulation& operator+=(const integration& in);
    int64_t countdofs() const { return dofm_->countdofs(); }
    // Builds the plan description in the reference's fragment order.
    std::shared_ptr<PlanStorage> buildplan();
    DofManager& dofmanager() { return *dofm_; }
    // Orientation folding: orientation batches whose bases are signed
    // renumberings of orientation 0 share one bucket (on by default).
    void set_orientation_folding(bool on) { fold_ = on; }

   private:
    struct Contribution {
        int integral_id;
        std::shared_ptr<RawField> doffield, tffield;
        int doffam = 0, tffam = 0;  // hcurl table family of the dof / tf (0: H1)
        std::vector<OpP> coeffs, dofs, tfs;
        int intphysreg, dofphysreg, tfphysreg, delta;
    };
    const Mesh* mesh_;
    std::shared_ptr<DofManager> dofm_;
    std::vector<Contribution> contribs_[2];  // [0] rhs, [1] K
    int nintegrals_ = 0;
    bool fold_ = true;
};

}  // namespace slx
