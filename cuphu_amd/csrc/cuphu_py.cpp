// cuphu_py.cpp -- pybind11 module _cuphu_ext (reference:
// src/cxx/cuphu/cuphu_py.cu).  Same functions and keyword arguments as the
// reference module: unwrap_arrays, build_costs, gpu_count, gpu_name,
// gpu_info; plus regrow_conncomp_arrays.  It only calls the C ABI of
// libcuphu_hip.so and releases the GIL for the whole GPU call.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <complex>
#include <functional>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "cuphu.h"

namespace py = pybind11;
using namespace pybind11::literals;

namespace {

CuPhuCostMode parse_cost_mode(const std::string &s) {
    if (s == "smooth") return CUPHU_COST_SMOOTH;
    if (s == "defo") return CUPHU_COST_DEFO;
    if (s == "topo") return CUPHU_COST_TOPO;
    throw py::value_error("cost must be 'smooth', 'defo', or 'topo'");
}

CuPhuInitMethod parse_init_method(const std::string &s) {
    if (s == "mst") return CUPHU_INIT_MST;
    if (s == "mcf") return CUPHU_INIT_MCF;
    if (s == "laplace") return CUPHU_INIT_LAPLACE;
    throw py::value_error("init must be 'mst', 'mcf', or 'laplace'");
}

CuPhuRampType parse_ramp_type(const py::object &obj) {
    if (obj.is_none()) return CUPHU_RAMP_NONE;
    std::string s = obj.cast<std::string>();
    if (s == "linear") return CUPHU_RAMP_LINEAR;
    if (s == "quadratic") return CUPHU_RAMP_QUADRATIC;
    if (s == "linear_range") return CUPHU_RAMP_LINEAR_RANGE;
    if (s == "linear_azimuth") return CUPHU_RAMP_LINEAR_AZIMUTH;
    if (s == "quadratic_range") return CUPHU_RAMP_QUADRATIC_RANGE;
    if (s == "quadratic_azimuth") return CUPHU_RAMP_QUADRATIC_AZIMUTH;
    throw py::value_error(
        "bridge_ramp_type must be one of: linear, quadratic, linear_range, linear_azimuth, quadratic_range, "
        "quadratic_azimuth, or None");
}

// Maps a C-ABI status to a Python exception (the reference raises
// RuntimeError for every failure, cuphu_py.cu:238-239).
void raise_status(int rc, const char *what) {
    if (rc >= 0) return;
    std::string msg = std::string(what) + " failed with code " + std::to_string(rc) + ": " + cuphu_last_error();
    if (rc == CUPHU_ERR_UNSUPPORTED) throw py::value_error(msg);
    if (rc == CUPHU_ERR_INVALID) throw py::value_error(msg);
    throw std::runtime_error(msg);
}

template <class S> void require_2d(const py::array_t<S, py::array::c_style> &a, const char *name) {
    if (a.ndim() != 2) throw py::value_error(std::string(name) + " must be 2-D");
}

CuPhuParams make_params(double nlooks, double costscale, int nshortcycle, int kperpdpsi, int kpardpsi, double defomax,
                        double min_conncomp_frac, long max_ncomps, long conncompthresh) {
    // cuphu_py.cu:85-109
    CuPhuParams p;
    cuphu_default_params(&p);
    p.nlooks = nlooks;
    p.ncorrlooks = nlooks;
    p.costscale = costscale;
    p.nshortcycle = nshortcycle;
    p.kperpdpsi = kperpdpsi;
    p.kpardpsi = kpardpsi;
    p.defomax = defomax;
    p.minconncompfrac = min_conncomp_frac;
    p.maxncomps = max_ncomps;
    p.conncompthresh = conncompthresh;
    return p;
}

py::tuple py_unwrap_arrays(py::array_t<std::complex<float>, py::array::c_style> igram,
                           py::array_t<float, py::array::c_style> corr, double nlooks, std::string cost,
                           std::string init, py::object mask_obj, py::object mag_obj, py::object orig_mask_obj,
                           double costscale, int nshortcycle, int kperpdpsi, int kpardpsi, double defomax,
                           double min_conncomp_frac, long max_ncomps, long conncompthresh, int ntilerow, int ntilecol,
                           int tile_rowovrlp, int tile_colovrlp, int tilecostthresh, int minregionsize, int nproc,
                           bool single_tile_reoptimize, bool laplace_neighbor_feedback,
                           int laplace_neighbor_feedback_feather, bool fix_cycle_spikes, bool bridge,
                           int bridge_radius, int bridge_min_num_pixel, int bridge_erosion_size,
                           int bridge_max_boundary_samples, py::object bridge_ramp_type_obj,
                           int bridge_ramp_max_num_sample, int gpu_id, py::object options) {
    require_2d(igram, "igram");
    require_2d(corr, "corr");
    int nrow = (int)igram.shape(0), ncol = (int)igram.shape(1);
    if (corr.shape(0) != nrow || corr.shape(1) != ncol) throw py::value_error("corr must have the same shape as igram");
    if (nlooks < 1.0) throw py::value_error("nlooks must be >= 1");
    (void)mag_obj;   // the Laplace path never reads the magnitude

    py::array_t<uint8_t, py::array::c_style> mask_arr, orig_arr;
    const unsigned char *mask_ptr = nullptr, *orig_ptr = nullptr;
    if (!mask_obj.is_none()) {
        mask_arr = mask_obj.cast<py::array_t<uint8_t, py::array::c_style>>();
        require_2d(mask_arr, "mask");
        if (mask_arr.shape(0) != nrow || mask_arr.shape(1) != ncol) throw py::value_error("mask shape mismatch");
        mask_ptr = mask_arr.data();
    }
    if (!orig_mask_obj.is_none()) {
        orig_arr = orig_mask_obj.cast<py::array_t<uint8_t, py::array::c_style>>();
        require_2d(orig_arr, "orig_mask");
        if (orig_arr.shape(0) != nrow || orig_arr.shape(1) != ncol) throw py::value_error("orig_mask shape mismatch");
        orig_ptr = orig_arr.data();
    }
    CuPhuParams params = make_params(nlooks, costscale, nshortcycle, kperpdpsi, kpardpsi, defomax, min_conncomp_frac,
                                     max_ncomps, conncompthresh);
    CuPhuTileParams tile;
    cuphu_default_tile_params(&tile);
    tile.ntilerow = ntilerow;
    tile.ntilecol = ntilecol;
    tile.rowovrlp = tile_rowovrlp;
    tile.colovrlp = tile_colovrlp;
    tile.tilecostthresh = tilecostthresh;
    tile.minregionsize = minregionsize;
    tile.nproc = nproc;
    tile.single_tile_reoptimize = single_tile_reoptimize ? 1 : 0;
    tile.laplace_neighbor_feedback = laplace_neighbor_feedback ? 1 : 0;
    tile.laplace_neighbor_feedback_feather = laplace_neighbor_feedback_feather;
    tile.fix_cycle_spikes = fix_cycle_spikes ? 1 : 0;
    CuPhuBridgeParams bp;
    cuphu_default_bridge_params(&bp);
    bp.enabled = bridge ? 1 : 0;
    bp.radius = bridge_radius;
    bp.min_num_pixel = bridge_min_num_pixel;
    bp.erosion_size = bridge_erosion_size;
    bp.max_boundary_samples = bridge_max_boundary_samples;
    bp.ramp_type = parse_ramp_type(bridge_ramp_type_obj);
    bp.ramp_max_num_sample = bridge_ramp_max_num_sample;

    CuPhuLaplaceOptions opt;
    cuphu_default_laplace_options(&opt);
    if (!options.is_none()) {
        py::dict d = options.cast<py::dict>();
        for (auto item : d) {
            std::string k = item.first.cast<std::string>();
            py::handle v = item.second;
            if (k == "n_irls") opt.n_irls = v.cast<int>();
            else if (k == "delta_max") opt.delta_max = v.cast<float>();
            else if (k == "delta_min") opt.delta_min = v.cast<float>();
            else if (k == "max_iter") opt.max_iter = v.cast<int>();
            else if (k == "tol") opt.tol = v.cast<double>();
            else if (k == "solver") {
                std::string s = v.cast<std::string>();
                if (s == "plain") opt.solver = CUPHU_SOLVER_PLAIN;
                else if (s == "batched") opt.solver = CUPHU_SOLVER_BATCHED;
                else if (s == "fused") opt.solver = CUPHU_SOLVER_FUSED;
                else throw py::value_error("solver must be 'plain', 'batched' or 'fused'");
            } else if (k == "precision") opt.precision = v.cast<int>();
            else if (k == "rr_interval") opt.rr_interval = v.cast<int>();
            else if (k == "use_graph") opt.use_graph = v.cast<bool>() ? 1 : 0;
            else if (k == "group_bytes_mb") opt.group_bytes_mb = v.cast<int>();
            else if (k == "verbose") opt.verbose = v.cast<bool>() ? 1 : 0;
            else if (k == "autotune") opt.autotune = v.cast<bool>() ? 1 : 0;
            else if (k == "stitch_integer") opt.stitch_integer = v.cast<bool>() ? 1 : 0;
            else if (k == "merge_labels") opt.merge_labels = v.cast<bool>() ? 1 : 0;
            else throw py::value_error("unknown laplace option '" + k + "'");
        }
    }
    CuPhuCostMode cm = parse_cost_mode(cost);
    CuPhuInitMethod im = parse_init_method(init);
    CuPhuResult result = {};
    std::vector<CuPhuSolveStats> stats((size_t)std::max(1, ntilerow * ntilecol));
    const float *ig = reinterpret_cast<const float *>(igram.data());
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = cuphu_unwrap_ex(nullptr, nullptr, ig, corr.data(), mask_ptr, nrow, ncol, cm, im, &params, &tile, &bp,
                             orig_ptr, gpu_id, &opt, stats.data(), &result);
    }
    raise_status(rc, "cuphu_unwrap");
    py::capsule free_unw(result.unw, [](void *p) { std::free(p); });
    py::capsule free_cc(result.conncomp, [](void *p) { std::free(p); });
    auto unw = py::array_t<float>({(py::ssize_t)nrow, (py::ssize_t)ncol},
                                  {(py::ssize_t)(ncol * sizeof(float)), (py::ssize_t)sizeof(float)}, result.unw, free_unw);
    auto cc = py::array_t<uint32_t>({(py::ssize_t)nrow, (py::ssize_t)ncol},
                                    {(py::ssize_t)(ncol * sizeof(uint32_t)), (py::ssize_t)sizeof(uint32_t)},
                                    result.conncomp, free_cc);
    return py::make_tuple(unw, cc);
}

py::array py_build_costs(py::array_t<std::complex<float>, py::array::c_style> igram,
                         py::array_t<float, py::array::c_style> corr, double nlooks, std::string cost, double costscale,
                         int nshortcycle, int kperpdpsi, int kpardpsi, double defomax, int gpu_id) {
    require_2d(igram, "igram");
    require_2d(corr, "corr");
    int nrow = (int)igram.shape(0), ncol = (int)igram.shape(1);
    if (corr.shape(0) != nrow || corr.shape(1) != ncol) throw py::value_error("corr must have the same shape as igram");
    // the reference splits the complex input into planes (cuphu_py.cu:278-285)
    std::vector<float> igr((size_t)nrow * ncol), igi((size_t)nrow * ncol);
    const std::complex<float> *src = igram.data();
    for (size_t k = 0; k < igr.size(); ++k) { igr[k] = src[k].real(); igi[k] = src[k].imag(); }
    CuPhuParams params = make_params(nlooks, costscale, nshortcycle, kperpdpsi, kpardpsi, defomax, 0.01, 32, 300);
    void *costs = nullptr;
    size_t esz = 0;
    CuPhuCostMode cm = parse_cost_mode(cost);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = cuphu_build_costs_gpu(igr.data(), igi.data(), corr.data(), nullptr, nullptr, nrow, ncol, cm, &params,
                                   gpu_id, &costs, &esz);
    }
    raise_status(rc, "cuphu_build_costs_gpu");
    size_t nb = ((size_t)(nrow - 1) * ncol + (size_t)nrow * (ncol - 1)) * esz;
    py::capsule free_costs(costs, [](void *p) { std::free(p); });
    return py::array_t<uint8_t>({(py::ssize_t)nb}, {(py::ssize_t)1}, (uint8_t *)costs, free_costs);
}

py::array py_regrow_conncomp(py::array_t<float, py::array::c_style> unw, py::array_t<float, py::array::c_style> corr,
                             py::object mask_obj, double cost_thresh, double min_conncomp_frac, long max_ncomps,
                             int gpu_id) {
    require_2d(unw, "unw");
    require_2d(corr, "corr");
    int nrow = (int)unw.shape(0), ncol = (int)unw.shape(1);
    if (corr.shape(0) != nrow || corr.shape(1) != ncol) throw py::value_error("corr shape mismatch");
    py::array_t<uint8_t, py::array::c_style> mask_arr;
    const unsigned char *mp = nullptr;
    if (!mask_obj.is_none()) {
        mask_arr = mask_obj.cast<py::array_t<uint8_t, py::array::c_style>>();
        if (mask_arr.ndim() != 2 || mask_arr.shape(0) != nrow || mask_arr.shape(1) != ncol)
            throw py::value_error("mask shape mismatch");
        mp = mask_arr.data();
    }
    py::array_t<uint32_t> out({(py::ssize_t)nrow, (py::ssize_t)ncol});
    uint32_t *op = out.mutable_data();
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = cuphu_regrow_conncomp(unw.data(), corr.data(), mp, nrow, ncol, cost_thresh, min_conncomp_frac, max_ncomps,
                                   gpu_id, op);
    }
    raise_status(rc, "cuphu_regrow_conncomp");
    return out;
}

int py_gpu_count() { return cuphu_gpu_count(); }

std::string py_gpu_name(int device_id) {
    char buf[256];
    raise_status(cuphu_gpu_name(device_id, buf, (int)sizeof(buf)), "gpu_name");
    return std::string(buf);
}

py::dict py_gpu_info(int device_id) {
    char name[256], arch[128];
    long long mem = 0;
    int cus = 0;
    raise_status(cuphu_gpu_name(device_id, name, (int)sizeof(name)), "gpu_info");
    raise_status(cuphu_gpu_info(device_id, &mem, &cus, arch, (int)sizeof(arch)), "gpu_info");
    py::dict d;
    d["name"] = std::string(name);
    d["total_memory"] = mem;
    d["sm_count"] = cus;
    d["compute_cap"] = std::string(arch);
    return d;
}

// Test-only bindings with the reference's names and arguments
// (cuphu_py.cu:373-431): run a post-pass on a copy of a caller-supplied
// unw array and return it.
py::array_t<float> post_pass_copy(const py::array_t<float, py::array::c_style> &unw, const py::object &mask_obj,
                                  const std::function<int(float *, const unsigned char *, int, int)> &fn,
                                  const char *what) {
    require_2d(unw, "unw");
    int nrow = (int)unw.shape(0), ncol = (int)unw.shape(1);
    py::array_t<uint8_t, py::array::c_style> m;
    const unsigned char *mp = nullptr;
    if (!mask_obj.is_none()) {
        m = mask_obj.cast<py::array_t<uint8_t, py::array::c_style>>();
        require_2d(m, "mask");
        if (m.shape(0) != nrow || m.shape(1) != ncol) throw py::value_error("mask must have the same shape as unw");
        mp = m.data();
    }
    py::array_t<float> out({(py::ssize_t)nrow, (py::ssize_t)ncol});
    std::memcpy(out.mutable_data(), unw.data(), (size_t)nrow * ncol * sizeof(float));
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = fn(out.mutable_data(), mp, nrow, ncol);
    }
    raise_status(rc, what);
    return out;
}

py::array_t<float> py_neighbor_feedback_test(py::array_t<float, py::array::c_style> unw, py::object mask, int ntilerow,
                                             int ntilecol, int row_ovrlp, int col_ovrlp, int feather_px) {
    return post_pass_copy(unw, mask, [&](float *u, const unsigned char *mp, int nr, int nc) {
        return cuphu_laplace_neighbor_feedback_test(u, mp, nr, nc, ntilerow, ntilecol, row_ovrlp, col_ovrlp, feather_px);
    }, "cuphu_laplace_neighbor_feedback_test");
}

py::array_t<float> py_fix_cycle_spikes_test(py::array_t<float, py::array::c_style> unw, py::object mask) {
    return post_pass_copy(unw, mask, [](float *u, const unsigned char *mp, int nr, int nc) {
        return cuphu_fix_cycle_spikes_test(u, mp, nr, nc);
    }, "cuphu_fix_cycle_spikes_test");
}

// Test-only binding with the reference's name and arguments
// (cuphu_py.cu:323-366): phase bridging on a copy of an unw array; returns
// (unw, conncomp) with conncomp starting at 0.
py::tuple py_bridge_apply_test(py::array_t<float, py::array::c_style> unw, py::object mask_obj, int bridge_radius,
                               int bridge_min_num_pixel, int bridge_erosion_size, int bridge_max_boundary_samples,
                               py::object bridge_ramp_type_obj, int bridge_ramp_max_num_sample, int gpu_id) {
    require_2d(unw, "unw");
    int nrow = (int)unw.shape(0), ncol = (int)unw.shape(1);
    py::array_t<uint8_t, py::array::c_style> m;
    const unsigned char *mp = nullptr;
    if (!mask_obj.is_none()) {
        m = mask_obj.cast<py::array_t<uint8_t, py::array::c_style>>();
        require_2d(m, "mask");
        if (m.shape(0) != nrow || m.shape(1) != ncol) throw py::value_error("mask must have the same shape as unw");
        mp = m.data();
    }
    CuPhuBridgeParams bp;
    cuphu_default_bridge_params(&bp);
    bp.enabled = 1;
    bp.radius = bridge_radius;
    bp.min_num_pixel = bridge_min_num_pixel;
    bp.erosion_size = bridge_erosion_size;
    bp.max_boundary_samples = bridge_max_boundary_samples;
    bp.ramp_type = parse_ramp_type(bridge_ramp_type_obj);
    bp.ramp_max_num_sample = bridge_ramp_max_num_sample;
    py::array_t<float> out({(py::ssize_t)nrow, (py::ssize_t)ncol});
    py::array_t<uint32_t> cc({(py::ssize_t)nrow, (py::ssize_t)ncol});
    std::memcpy(out.mutable_data(), unw.data(), (size_t)nrow * ncol * sizeof(float));
    std::memset(cc.mutable_data(), 0, (size_t)nrow * ncol * sizeof(uint32_t));
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = cuphu_bridge_apply_test(out.mutable_data(), cc.mutable_data(), nrow, ncol, &bp, mp, gpu_id);
    }
    raise_status(rc, "cuphu_bridge_apply_test");
    return py::make_tuple(out, cc);
}

}  // namespace

PYBIND11_MODULE(_cuphu_ext, m) {
    m.doc() = "MI355X (gfx950) Laplace PCG/IRLS phase unwrapping (HIP extension)";
    m.def("unwrap_arrays", &py_unwrap_arrays, "igram"_a, "corr"_a, "nlooks"_a, "cost"_a = "smooth", "init"_a = "mcf",
          "mask"_a = py::none(), "mag"_a = py::none(), "orig_mask"_a = py::none(), "costscale"_a = 100.0,
          "nshortcycle"_a = 200, "kperpdpsi"_a = 7, "kpardpsi"_a = 7, "defomax"_a = 1.2, "min_conncomp_frac"_a = 0.01,
          "max_ncomps"_a = 32L, "conncompthresh"_a = 300L, "ntilerow"_a = 1, "ntilecol"_a = 1, "tile_rowovrlp"_a = 0,
          "tile_colovrlp"_a = 0, "tilecostthresh"_a = 500, "minregionsize"_a = 100, "nproc"_a = 1,
          "single_tile_reoptimize"_a = false, "laplace_neighbor_feedback"_a = false,
          "laplace_neighbor_feedback_feather"_a = 200, "fix_cycle_spikes"_a = false, "bridge"_a = false,
          "bridge_radius"_a = 500, "bridge_min_num_pixel"_a = 14, "bridge_erosion_size"_a = 2,
          "bridge_max_boundary_samples"_a = 4096, "bridge_ramp_type"_a = py::none(),
          "bridge_ramp_max_num_sample"_a = 1000000, "gpu_id"_a = 0, "laplace_options"_a = py::none(),
          "Unwrap an interferogram with the GPU Laplace PCG/IRLS path; returns (unw float32, conncomp uint32).");
    m.def("build_costs", &py_build_costs, "igram"_a, "corr"_a, "nlooks"_a, "cost"_a = "smooth", "costscale"_a = 100.0,
          "nshortcycle"_a = 200, "kperpdpsi"_a = 7, "kpardpsi"_a = 7, "defomax"_a = 1.2, "gpu_id"_a = 0,
          "Build GPU cost arrays and return them as a raw byte buffer.");
    m.def("regrow_conncomp_arrays", &py_regrow_conncomp, "unw"_a, "corr"_a, "mask"_a = py::none(),
          "cost_thresh"_a = 0.1, "min_conncomp_frac"_a = 0.01, "max_ncomps"_a = 32L, "gpu_id"_a = 0,
          "Connected components of an unwrapped phase on the GPU.");
    m.def("_laplace_neighbor_feedback_test", &py_neighbor_feedback_test, "unw"_a, "mask"_a = py::none(), "ntilerow"_a,
          "ntilecol"_a, "row_ovrlp"_a, "col_ovrlp"_a, "feather_px"_a = 200,
          "Test-only: the Laplace tile-seam neighbour feedback on a copy of an already-stitched unw array.");
    m.def("_bridge_apply_test", &py_bridge_apply_test, "unw"_a, "mask"_a = py::none(), "bridge_radius"_a = 500,
          "bridge_min_num_pixel"_a = 14, "bridge_erosion_size"_a = 2, "bridge_max_boundary_samples"_a = 4096,
          "bridge_ramp_type"_a = py::none(), "bridge_ramp_max_num_sample"_a = 1000000, "gpu_id"_a = 0,
          "Test-only: phase bridging on a copy of an unw array; returns (unw, conncomp).");
    m.def("_fix_cycle_spikes_test", &py_fix_cycle_spikes_test, "unw"_a, "mask"_a = py::none(),
          "Test-only: the isolated row/column whole-cycle spike correction on a copy of an unw array.");
    m.def("gpu_count", &py_gpu_count, "Return the number of GPUs available.");
    m.def("gpu_name", &py_gpu_name, "device_id"_a = 0, "Return the name of the device.");
    m.def("gpu_info", &py_gpu_info, "device_id"_a = 0, "Return a dict of device properties.");
}
