// survey.cpp -- see survey.hpp.
#include "survey.hpp"

#include <limits>
#include <cstdio>
#include <cstring>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iostream>
#include <stdexcept>

namespace fwi {

void check(int rc, const char *what)
{
    if (rc != FWI_OK)
        throw std::runtime_error(std::string(what) + ": " + fwi_last_error());
}

namespace {

std::string fname_suffix(float fmax, int nz, int nx, float dh)
{
    return std::to_string(int(fmax)) + "Hz_" + std::to_string(nz) + "x" + std::to_string(nx) +
           "_" + std::to_string((int)(dh)) + "m.bin";
}

}  // namespace

Modeling::~Modeling()
{
    if (ctx) fwi_destroy(ctx);
}

void Modeling::set_main_parameters()
{
    nx = param_int("x_samples", parameters);
    nz = param_int("z_samples", parameters);
    dh = param_float("model_spacing", parameters);
    nt = param_int("time_samples", parameters);
    dt = param_float("time_spacing", parameters);
    fmax = param_float("max_frequency", parameters);
}

void Modeling::set_wavelet()
{
    wavelet.assign(nt, 0.f);
    int32_t tl = 0;
    check(fwi_ricker_wavelet(nt, dt, fmax, wavelet.data(), &tl), "wavelet");
    tlag = tl;
}

void Modeling::set_properties()
{
    vp.assign((size_t)nx * nz, 0.f);
    import_binary_float(catch_parameter("model_file", parameters), vp.data(), (long)nx * nz);
    vmin = *std::min_element(vp.begin(), vp.end());
    vmax = *std::max_element(vp.begin(), vp.end());
}

// Kind of random boundary: `rbc_kind = gaussian` (the default: the
// reference's Gaussian noise and Poisson-disk blobs of radius up to
// *_rbc_ratio, seeded) or `uniform` (a uniform perturbation of amplitude
// *_rbc_varVp + *_rbc_varVp_rel times the edge speed).
int rbc_kind_of(const std::string &parameters)
{
    const std::string k = catch_parameter("rbc_kind", parameters, "gaussian");
    if (k == "gaussian") return FWI_RBC_GAUSSIAN;
    if (k == "uniform") return FWI_RBC_UNIFORM;
    throw std::runtime_error("Error: rbc_kind must be gaussian or uniform, not '" + k + "'");
}

void Modeling::create_engine(int rbc_nb, float varVp, float varVp_rel, uint64_t seed, int kind,
                             float ratio)
{
    geometry.set_parameters(parameters);
    max_batch = param_int("gpu_max_batch", parameters, std::min(geometry.nsrc, 256));
    fwi_config cfg{};
    cfg.nz = nz;
    cfg.nx = nx;
    cfg.dh = dh;
    cfg.nt = nt;
    cfg.dt = dt;
    cfg.abc_nb = abc_nb;
    cfg.abc_factor = abc_factor;
    cfg.rbc_nb = rbc_nb;
    cfg.rbc_varVp = varVp;
    cfg.rbc_varVp_rel = varVp_rel;
    cfg.rbc_seed = seed;
    cfg.rbc_kind = kind;
    cfg.rbc_ratio = ratio;
    cfg.max_batch = std::max(1, max_batch);
    check(fwi_create(&cfg, &ctx), "create");
    check(fwi_set_wavelet(ctx, wavelet.data(), tlag), "set_wavelet");
    check(fwi_set_geometry(ctx, geometry.nsrc, geometry.xsrc.data(), geometry.zsrc.data(),
                           geometry.nrec, geometry.xrec.data(), geometry.zrec.data()),
          "set_geometry");
    check(fwi_set_model(ctx, vp.data()), "set_model");
    // launch shapes timed on this grid and batch (about a second at
    // production sizes); results do not depend on them
    if (param_bool("gpu_autotune", parameters, true)) {
        double fwd_us = 0.0, rev_us = 0.0;
        check(fwi_autotune(ctx, std::max(1, std::min(max_batch, geometry.nsrc)), 8, &fwd_us,
                           &rev_us),
              "autotune");
    }
}

void Modeling::set_parameters()
{
    title = "Seismic Modeling";
    set_main_parameters();
    const float abc_length = param_float("abc_length", parameters);
    abc_factor = param_float("abc_factor", parameters);
    abc_nb = (int)(abc_length / dh) + 1;
    set_wavelet();
    set_properties();
    data_folder = catch_parameter("mod_output_folder", parameters, "");
    create_engine(0, 0.f, 0.f, 0);
}

void Modeling::show_information() const
{
    std::cout << std::string(80, '-') << "\n" << title << "\n" << std::string(80, '-') << "\n";
    std::cout << "Model dimensions: (z = " << (nz - 1) * dh << ", x = " << (nx - 1) * dh
              << ") m\n";
    std::cout << "Shots: " << geometry.nsrc << ", receivers: " << geometry.nrec
              << ", batch: " << max_batch << "\n";
}

void Modeling::forward_solver()
{
    // no zero fill: every sample is written by the call (1 GB at 64 shots of
    // 1001 receivers x 4001 samples)
    seismogram.reset(new float[(size_t)geometry.nsrc * geometry.nrec * nt]);
    check(fwi_modeling(ctx, FWI_BOUNDARY_ABC, 0, geometry.nsrc, seismogram.get()), "modeling");
}

void Modeling::get_seismogram()
{
    const long per = (long)geometry.nrec * nt;
    for (int s = 0; s < geometry.nsrc; ++s) {
        std::string file = data_folder + "seismogram_nt" + std::to_string(nt) + "_nr" +
                           std::to_string(geometry.nrec) + "_" + std::to_string(int(1e6f * dt)) +
                           "us_shot_" + std::to_string(s + 1) + ".bin";
        export_binary_float(file, seismogram.get() + s * per, per);
    }
}

// ------------------------------------------------------------------ Migration

void Migration::set_parameters()
{
    title = "Reverse Time Migration";
    set_main_parameters();
    rbc_varVp = param_float("mig_rbc_varVp", parameters);
    rbc_varVp_rel = param_float("mig_rbc_varVp_rel", parameters, 0.f);
    rbc_ratio = param_float("mig_rbc_ratio", parameters, 0.f);
    rbc_kind = rbc_kind_of(parameters);
    const float rbc_length = param_float("mig_rbc_length", parameters);
    rbc_nb = (int)(rbc_length / dh) + 1;
    const uint64_t seed = param_u64("rbc_seed", parameters, 20240u);
    set_wavelet();
    set_properties();
    input_folder = catch_parameter("mig_input_folder", parameters, "");
    input_prefix = catch_parameter("mig_input_prefix", parameters);
    output_folder = catch_parameter("mig_output_folder", parameters, "");
    abc_nb = 0;
    create_engine(rbc_nb, rbc_varVp, rbc_varVp_rel, seed, rbc_kind, rbc_ratio);
    const long per = (long)geometry.nrec * nt;
    std::vector<float> data(per);
    for (int s = 0; s < geometry.nsrc; ++s) {   // migration.cu:83-88, once per shot
        import_binary_float(input_folder + input_prefix + std::to_string(s + 1) + ".bin",
                            data.data(), per);
        check(fwi_set_observed(ctx, s, 1, data.data()), "set_observed");
    }
    check(fwi_reset(ctx, FWI_RESET_IMAGE | FWI_RESET_ILLUM), "reset");
}

void Migration::run_migration()
{
    check(fwi_migration(ctx, 0, geometry.nsrc), "migration");
}

void Migration::export_seismic()
{
    std::vector<float> image((size_t)nx * nz);
    check(fwi_get_image(ctx, image.data(), nullptr, nullptr), "get_image");
    export_binary_float(output_folder + "RTM_section_" + fname_suffix(fmax, nz, nx, dh),
                        image.data(), (long)nx * nz);
}

// ------------------------------------------------------------------ Inversion

Inversion::~Inversion()
{
    if (net) fwi_nn_destroy(net);
}

// inv_trial_direction = none | plain | network, inv_trial_network = file of
// the network parameters (fwi2d_nn.h layout, raw float32),
// inv_trial_multiples = k1,k2,... (1 to 4 multiples of the plain step).
void Inversion::set_trial()
{
    const std::string kind = catch_parameter("inv_trial_direction", parameters, "none");
    if (kind == "none") return;
    if (kind == "plain") trial = FWI_TRIAL_PLAIN;
    else if (kind == "network") trial = FWI_TRIAL_NETWORK;
    else throw std::runtime_error("Error: inv_trial_direction must be none, plain or network, "
                                  "not '" + kind + "'");
    for (const std::string &k : split(catch_parameter("inv_trial_multiples", parameters, "2,4"), ','))
        multiples.push_back(std::stod(k));
    if (multiples.empty() || multiples.size() > 4)
        throw std::runtime_error("Error: inv_trial_multiples takes 1 to 4 values");
    if (trial != FWI_TRIAL_NETWORK) return;
    const int H = nx + (4 - nx % 4) % 4, W = nz + (4 - nz % 4) % 4;
    check(fwi_nn_create(1, H, W, 0, &net), "network");
    int64_t np = 0;
    check(fwi_nn_num_params(net, &np), "network size");
    std::vector<float> p((size_t)np);
    const std::string file = catch_parameter("inv_trial_network", parameters);
    import_binary_float(file, p.data(), (long)np);
    check(fwi_nn_set_params(net, p.data()), "network parameters");
    raw_grad.resize((size_t)nx * nz);
    raw_illum.resize((size_t)nx * nz);
}

void Inversion::set_parameters()
{
    title = "Full Waveform Inversion";
    set_main_parameters();
    const float abc_length = param_float("abc_length", parameters);
    abc_factor = param_float("abc_factor", parameters);
    abc_nb = (int)(abc_length / dh) + 1;
    rbc_varVp = param_float("inv_rbc_varVp", parameters);
    rbc_varVp_rel = param_float("inv_rbc_varVp_rel", parameters, 0.f);
    rbc_ratio = param_float("inv_rbc_ratio", parameters, 0.f);
    rbc_kind = rbc_kind_of(parameters);
    const float rbc_length = param_float("inv_rbc_length", parameters);
    rbc_nb = (int)(rbc_length / dh) + 1;
    const uint64_t seed = param_u64("rbc_seed", parameters, 20240u);
    zmask = param_float("depth_mask", parameters);
    max_iteration = param_int("max_iteration", parameters);
    input_folder = catch_parameter("inv_input_folder", parameters, "");
    input_prefix = catch_parameter("inv_input_prefix", parameters);
    output_folder = catch_parameter("inv_output_folder", parameters, "");
    residuo_folder = catch_parameter("inv_residuo_folder", parameters, "");
    checkpoint_file = catch_parameter("checkpoint_file", parameters, "");
    resume = param_bool("resume", parameters, false);
    if (resume && checkpoint_file.empty())
        throw std::runtime_error("Error: resume = true needs a checkpoint_file");
    set_wavelet();
    set_properties();
    vmin0 = vmin;
    vmax0 = vmax;
    set_trial();
    create_engine(rbc_nb, rbc_varVp, rbc_varVp_rel, seed, rbc_kind, rbc_ratio);
    // observed data stay resident on the device (read once, not every pass)
    const long per = (long)geometry.nrec * nt;
    std::vector<float> data(per);
    for (int s = 0; s < geometry.nsrc; ++s) {
        import_binary_float(input_folder + input_prefix + std::to_string(s + 1) + ".bin",
                            data.data(), per);
        check(fwi_set_observed(ctx, s, 1, data.data()), "set_observed");
    }
    model.resize(vp.size());
    for (size_t k = 0; k < vp.size(); ++k) model[k] = 1.0f / (vp[k] * vp[k]);
    gradient.assign(vp.size(), 0.f);
    // the illumination accumulates over every gradient pass (inversion.cu:51, 441)
    check(fwi_reset(ctx, FWI_RESET_GRADIENT | FWI_RESET_ILLUM), "reset");
    if (resume && load_checkpoint())
        std::cout << "resumed from " << checkpoint_file << " at iteration " << iteration << "\n";
}

namespace {
const char kCheckpointMagic[8] = {'F', 'W', 'I', '2', 'D', 'C', 'K', '1'};
}

void Inversion::save_checkpoint() const
{
    if (checkpoint_file.empty()) return;
    int32_t nzz = 0, nxx = 0, nb = 0;
    check(fwi_grid_shape(ctx, FWI_BOUNDARY_RBC, &nzz, &nxx, &nb), "grid_shape");
    std::vector<float> illum((size_t)nzz * nxx);
    check(fwi_get_accumulator(ctx, FWI_RESET_ILLUM, illum.data()), "get illumination");
    const std::string tmp = checkpoint_file + ".tmp";
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f.is_open()) throw std::runtime_error("Error: " + tmp + " could not be opened!");
        const int32_t head[6] = {nz, nx, nzz, nxx, iteration, (int32_t)residuo.size()};
        f.write(kCheckpointMagic, 8);
        f.write(reinterpret_cast<const char *>(head), sizeof head);
        f.write(reinterpret_cast<const char *>(&vmin0), 4);
        f.write(reinterpret_cast<const char *>(&vmax0), 4);
        f.write(reinterpret_cast<const char *>(residuo.data()), residuo.size() * 8);
        f.write(reinterpret_cast<const char *>(model.data()), model.size() * 4);
        f.write(reinterpret_cast<const char *>(illum.data()), illum.size() * 4);
        if (!f) throw std::runtime_error("Error: writing " + tmp + " failed");
    }
    if (std::rename(tmp.c_str(), checkpoint_file.c_str()) != 0)
        throw std::runtime_error("Error: could not replace " + checkpoint_file);
}

bool Inversion::load_checkpoint()
{
    std::ifstream f(checkpoint_file, std::ios::binary);
    if (!f.is_open()) return false;               // nothing to resume: start from scratch
    char magic[8];
    int32_t head[6];
    f.read(magic, 8);
    f.read(reinterpret_cast<char *>(head), sizeof head);
    if (!f || std::memcmp(magic, kCheckpointMagic, 8) != 0)
        throw std::runtime_error("Error: " + checkpoint_file + " is not a checkpoint");
    int32_t nzz = 0, nxx = 0, nb = 0;
    check(fwi_grid_shape(ctx, FWI_BOUNDARY_RBC, &nzz, &nxx, &nb), "grid_shape");
    if (head[0] != nz || head[1] != nx || head[2] != nzz || head[3] != nxx)
        throw std::runtime_error("Error: " + checkpoint_file + " belongs to another grid");
    if (head[4] < 0 || head[5] != head[4])
        throw std::runtime_error("Error: " + checkpoint_file + " is inconsistent");
    float v0[2];
    f.read(reinterpret_cast<char *>(v0), 8);
    std::vector<double> res((size_t)head[5]);
    f.read(reinterpret_cast<char *>(res.data()), res.size() * 8);
    std::vector<float> m(model.size()), illum((size_t)nzz * nxx);
    f.read(reinterpret_cast<char *>(m.data()), m.size() * 4);
    f.read(reinterpret_cast<char *>(illum.data()), illum.size() * 4);
    if (!f) throw std::runtime_error("Error: " + checkpoint_file + " is truncated");
    vmin0 = v0[0];
    vmax0 = v0[1];
    residuo = res;
    iteration = head[4];
    model = m;
    for (size_t k = 0; k < model.size(); ++k) vp[k] = 1.0f / std::sqrt(model[k]);
    check(fwi_set_model(ctx, vp.data()), "set_model (resume)");
    check(fwi_set_accumulator(ctx, FWI_RESET_ILLUM, illum.data()), "set illumination");
    return true;
}

void Inversion::compute_gradient()
{
    sum_res = 0.0;
    check(fwi_reset(ctx, FWI_RESET_GRADIENT), "reset");
    if (iteration < max_iteration) {
        check(fwi_gradient(ctx, 0, geometry.nsrc, (uint32_t)iteration, &sum_res), "gradient");
        double gmax = 0.0;
        const bool raw = trial == FWI_TRIAL_NETWORK;   // the network reads the raw fields
        check(fwi_get_gradient(ctx, zmask, gradient.data(), raw ? raw_grad.data() : nullptr,
                               raw ? raw_illum.data() : nullptr, &gmax),
              "get_gradient");
        std::cout << "iteration " << iteration << ": gradient pass done, max|g| = " << gmax
                  << "\n";
    } else {
        // the final pass only measures the residual (inversion.cu:117-136)
        check(fwi_misfit(ctx, 0, geometry.nsrc, &sum_res), "misfit");
    }
}

void Inversion::check_convergence()
{
    ++iteration;
    residuo.push_back(std::sqrt(sum_res));
    converged = iteration > max_iteration;
    std::cout << "residuo[" << residuo.size() - 1 << "] = " << residuo.back() << "\n";
}

double Inversion::min_slowness2() const
{
    // 1/v^2 of the fastest velocity the propagator accepts (with a margin)
    const double vmax_stable = 0.999 * 0.5546 * (double)dh / (double)dt;
    return 1.0 / (vmax_stable * vmax_stable);
}

void Inversion::optimization()
{
    // trial steps 0.02 and 0.05 of 1/vmin^2 - 1/vmax^2 of the initial model
    // (inversion.cu:288-301)
    const double dm = 1.0 / ((double)vmin0 * vmin0) - 1.0 / ((double)vmax0 * vmax0);
    fwi_update_opts o{};
    o.alpha1 = 0.02 * dm;
    o.alpha2 = 0.05 * dm;
    o.m_min = min_slowness2();
    o.zmask = zmask;
    o.trial = trial;
    o.net = net;
    o.nmultiples = (int32_t)multiples.size();
    for (size_t k = 0; k < multiples.size(); ++k) o.multiples[k] = multiples[k];
    check(fwi_update_model(ctx, &o, residuo.back(), gradient.data(),
                           raw_grad.empty() ? nullptr : raw_grad.data(),
                           raw_illum.empty() ? nullptr : raw_illum.data(), model.data(), &last),
          "update");
    static const char *names[3] = {"parabola vertex", "best trial step", "no decrease, no step"};
    std::cout << "line search: f = " << last.f[0] << ", " << last.f[1] << ", " << last.f[2]
              << " -> step " << last.step << " (" << names[last.rule] << ")\n";
    if (trial != FWI_TRIAL_NONE && last.step > 0.0) {
        const auto prec = std::cout.precision(12);
        std::cout << "trial direction: misfit " << last.trial_misfit << " at " << last.trial_multiple
                  << " x step against " << last.plain_misfit << " for the plain step -> "
                  << (last.taken ? "trial taken" : "plain step") << "\n";
        std::cout.precision(prec);
    }
}

void Inversion::update_model()
{
    // fwi_update_model has updated m and left the model in the engine
    for (size_t k = 0; k < model.size(); ++k) vp[k] = 1.0f / std::sqrt(model[k]);
    export_binary_float(output_folder + "model_FWI_iteration_" + std::to_string(iteration) + "_" +
                            fname_suffix(fmax, nz, nx, dh),
                        vp.data(), (long)vp.size());
}

void Inversion::export_convergence()
{
    const std::string path =
        residuo_folder + "convergence_" + std::to_string(max_iteration) + "_iterations.txt";
    std::ofstream out(path);
    if (!out.is_open()) throw std::runtime_error("Error: " + path + " could not be opened!");
    for (double r : residuo) out << (float)r << "\n";
    std::cout << "Text file " << path << " was successfully written." << std::endl;
}

void Inversion::export_final_model()
{
    export_binary_float(output_folder + "final_model_FWI_" + fname_suffix(fmax, nz, nx, dh),
                        vp.data(), (long)vp.size());
}

}  // namespace fwi
