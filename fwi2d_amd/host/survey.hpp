// survey.hpp -- the workflow classes of the drop-in executables, with the
// reference's method names (src/modeling/modeling.cuh:18-87,
// src/migration/migration.cuh:6-38, src/inversion/inversion.cuh:6-83).
// All wavefield work runs on the GPU engine behind the C ABI (fwi2d.h); every
// shot of a pass runs as one batch instead of one shot at a time.
#pragma once

#include <string>
#include <memory>
#include <vector>

#include "admin.hpp"
#include "fwi2d.h"
#include "fwi2d_update.h"

namespace fwi {

// Throws std::runtime_error with the engine's message when rc != FWI_OK.
void check(int rc, const char *what);

class Modeling {
public:
    virtual ~Modeling();
    std::string parameters;
    Geometry geometry;

    void set_parameters();        // modeling.cu:3-15
    void forward_solver();        // all shots; modeling.cu:286-299
    void get_seismogram();        // one file per shot; modeling.cu:279-284
    void show_information() const;

protected:
    int nx = 0, nz = 0, nt = 0, tlag = 0, max_batch = 1;
    float dh = 0, dt = 0, fmax = 0;
    int abc_nb = 0;
    float abc_factor = 0;
    float vmin = 0, vmax = 0;
    std::vector<float> vp;          // interior model [nx][nz]
    std::vector<float> wavelet;
    std::unique_ptr<float[]> seismogram;  // [nsrc][nrec][nt], written whole by fwi_modeling
    std::string data_folder, title;
    fwi_ctx *ctx = nullptr;

    void set_main_parameters();     // modeling.cu:17-30
    void set_wavelet();             // modeling.cu:32-90
    void set_properties();          // modeling.cu:152-188
    void create_engine(int rbc_nb, float varVp, float varVp_rel, uint64_t seed,
                       int kind = FWI_RBC_UNIFORM, float ratio = 0.f);
};

class Migration : public Modeling {
public:
    void set_parameters();          // migration.cu:3-36
    void run_migration();           // forward_propagation + backward_propagation, all shots
    void export_seismic();          // migration.cu:90-115

private:
    float rbc_varVp = 0, rbc_varVp_rel = 0, rbc_ratio = 0;
    int rbc_nb = 0, rbc_kind = FWI_RBC_GAUSSIAN;
    std::string input_folder, input_prefix, output_folder;
};

class Inversion : public Modeling {
public:
    ~Inversion() override;
    bool converged = false;
    void set_parameters();          // inversion.cu:3-53
    void compute_gradient();        // inversion.cu:111-166
    void check_convergence();       // inversion.cu:275-284
    // The line search, the step and m <- m - step g (inversion.cu:286-342,
    // 344-349) in fwi_update_model; with inv_trial_direction = plain or
    // network it also tries the trial direction and keeps it only when its
    // forward-modeled misfit is below the plain update's.
    void optimization();
    void update_model();            // model file of the iteration (inversion.cu:350-358)
    void export_convergence();      // inversion.cu:360-372
    void export_final_model();      // inversion.cu:374-379
    // Checkpoint / resume (no reference counterpart): after every model
    // update the state that the next iteration needs -- iteration, residual
    // history, model, the initial model's vmin / vmax and the accumulated
    // illumination -- goes to `checkpoint_file`; with `resume = true` an
    // existing checkpoint is loaded before the first iteration.
    void save_checkpoint() const;
    bool load_checkpoint();

    int iteration = 0;
    int max_iteration = 0;
    std::vector<double> residuo;
    std::vector<float> model;       // m = 1/v^2 (inversion.cu:101-109)
    std::vector<float> gradient;    // search direction

private:
    float zmask = 0;
    float rbc_varVp = 0, rbc_varVp_rel = 0, rbc_ratio = 0;
    int rbc_nb = 0, rbc_kind = FWI_RBC_GAUSSIAN;
    double sum_res = 0;
    float vmin0 = 0, vmax0 = 0;     // extremes of the initial model (dm of the line search)
    std::string input_folder, input_prefix, output_folder, residuo_folder;
    std::string checkpoint_file;
    bool resume = false;
    // trial direction (no reference counterpart): FWI_TRIAL_NONE, _PLAIN or
    // _NETWORK, the network and its raw inputs, the multiples of the step
    int trial = FWI_TRIAL_NONE;
    fwi_nn *net = nullptr;
    std::vector<double> multiples;
    std::vector<float> raw_grad, raw_illum;
    fwi_update_info last{};         // record of the last update
    double min_slowness2() const;
    void set_trial();
};

}  // namespace fwi
