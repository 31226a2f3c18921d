// simulation.hpp -- Simulation : Model, the drop-in for the reference's
// driver class (src/Simulation/simulation.hpp:49-166).  Every method keeps the
// reference's name, arguments and return convention; the bodies call the
// gfx950 library through its C-ABI (include/magritte_hip.h).  The device keeps
// the model resident; results are copied back into the Model members in place
// (radiation.J, lambda, Jlin/Jeff/Jdif, population and its history) so that
// callers see the same fields as with the reference.
#pragma once

#include <string>

#include "../../../include/magritte_hip.h"
#include "image.hpp"
#include "model.hpp"

namespace magritte {

enum SpecDiscSetting { None, LineSet, ImageSet };

struct Simulation : public Model {
    Double1 error_max, error_mean;      // per iteration and species (sim_lines.tpp:137-138)
    Double1 timings;                    // radiation-field kernel time per call, ms
    Long1 nrpairs, depths;              // ray pairs solved, sum of their depths
    SpecDiscSetting specDiscSetting = None;

    int device = 0;                      // HIP device of the context
    int precision = MGR_FP32;            // of compute_radiation_field
    int image_precision = MGR_FP64;      // of compute_and_write_image (the reference's FP64)
    int convergence_criterion = MGR_CRIT_FRACTION;

    Simulation() = default;
    Simulation(const Simulation &) = delete;
    Simulation &operator=(const Simulation &) = delete;
    ~Simulation() override;

    int compute_spectral_discretisation();
    int compute_spectral_discretisation_image(const double width);
    int compute_boundary_intensities();
    int compute_boundary_intensities(const Double1 &temperatures);
    int compute_LTE_level_populations();

    int compute_radiation_field();        // FP32 solve (or `precision`)
    int compute_radiation_field_cpu();    // FP64 solve on the GPU (the reference's FP64 path)
    int compute_radiation_field_gpu();    // FP32 solve
    int gpu_compute_radiation_field(size_t nraypairs, size_t gpuBlockSize, size_t gpuNumBlocks,
                                    double inverse_dtau_max);
    int cpu_compute_radiation_field(size_t nraypairs, size_t gpuBlockSize, size_t gpuNumBlocks,
                                    double inverse_dtau_max);
    int gpu_get_device_properties();

    void compute_Jeff();
    void compute_level_populations_from_stateq();
    long compute_level_populations(const Io &io);
    long compute_level_populations(const Io &io, bool use_Ng_acceleration, long max_niterations);

    int compute_and_write_image(const Io &io, const long r);

    Long2 get_npoints_on_rays_comoving();
    Long2 get_npoints_on_rays_rest();

    // drop the device context (after the geometry or thermodynamics changed)
    void reset_device();

  private:
    mgr_ctx *ctx_ = nullptr;
    void ensure_context();
    void check(int rc, const char *what);
    void push_populations();
    void pull_populations();
    void pull_radiation();
    void pull_jeff();
    void pull_boundary_intensities();
    void pull_spectral();
    void pull_convergence();
    size_t nlines_total() const;
    size_t first_line(size_t l) const;
    bool use_scattering() const;
    double pop_prec() const { return parameters.pop_prec_.is_set ? parameters.pop_prec() : 1.0e-6; }
    int run_radiation_field(int prec);
};

}  // namespace magritte
