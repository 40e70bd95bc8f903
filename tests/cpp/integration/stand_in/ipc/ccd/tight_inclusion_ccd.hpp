// Stand-in: the reference's narrow-phase CCD base and Tight-Inclusion
// parameters (narrow_phase_ccd.hpp, tight_inclusion_ccd.hpp).
#pragma once

namespace ipc {

class NarrowPhaseCCD {
public:
    explicit NarrowPhaseCCD(double rescaling = 0.8) : conservative_rescaling(rescaling) { }
    virtual ~NarrowPhaseCCD() = default;
    double conservative_rescaling;
};

class TightInclusionCCD : public NarrowPhaseCCD {
public:
    static constexpr double DEFAULT_TOLERANCE = 1e-6;
    static constexpr long DEFAULT_MAX_ITERATIONS = 10'000'000L;
    static constexpr double DEFAULT_CONSERVATIVE_RESCALING = 0.8;
    explicit TightInclusionCCD(double tol = DEFAULT_TOLERANCE,
                               long max_iter = DEFAULT_MAX_ITERATIONS,
                               double rescaling = DEFAULT_CONSERVATIVE_RESCALING)
        : NarrowPhaseCCD(rescaling), tolerance(tol), max_iterations(max_iter)
    {
    }
    double tolerance;
    long max_iterations;
};

} // namespace ipc
