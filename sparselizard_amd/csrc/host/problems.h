// Synthetic workloads: meshes, fields and formulations of the benchmark
// configurations, generated with a portable PRNG so that every machine
// builds bit-identical inputs.
#pragma once
#include <map>
#include <memory>
#include <string>
#include "fem.h"

namespace slx {

// SplitMix64-seeded xoshiro256** with a Box-Muller normal sampler.
class Rng {
   public:
    explicit Rng(uint64_t seed);
    uint64_t next();
    double uniform();                       // [0, 1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    int uniform_int(int lo, int hi);        // inclusive
    double normal();

   private:
    uint64_t s_[4];
    bool has_spare_ = false;
    double spare_ = 0;
};

struct Problem {
    std::string kind;
    std::map<std::string, std::string> spec;
    std::shared_ptr<Mesh> mesh;
    std::vector<field> fields;
    std::vector<parameter> params;
    std::shared_ptr<formulation> form;
    std::shared_ptr<PlanStorage> plan;
    std::string describe;
};

std::map<std::string, std::string> parse_spec(const std::string& spec);
// hip_stream: the meshes are finalized on the device on that stream (null or
// the spec option meshpipe=host: on the host).
std::shared_ptr<Problem> make_problem(const std::string& kind, const std::string& spec, void* hip_stream = nullptr);

}  // namespace slx
