// Exercises the C++ interface (include/sie_mom/*.hpp) the way the reference's
// in-process callers do. Modes:
//   validate <tmpdir>          host-only checks (no GPU needed): option and
//                              source validation, OBJ loader, binary formats.
//   solve <dir> <freq> <rule>  GPU solves on an icosphere (level 1); writes
//                              currents for comparison with the oracle.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "sie_mom/gpu/dense_lu_solver.hpp"
#include "sie_mom/solvers/dielectric/dielectric_solver.hpp"
#include "sie_mom/mesh/circular_ogive.hpp"
#include "sie_mom/mesh/obj_mesh_loader.hpp"
#include "sie_mom/solvers/surface_current_fields.hpp"
#include "sie_mom/gpu/gmres_solver.hpp"
#include "sie_mom/solvers/gstc_sie.hpp"
#include "sie_mom/gstc_utils.hpp"
#include "sie_mom/solvers/pec_solver.hpp"

using namespace sie_mom;

static int failures = 0;
#define CHECK(cond)                                                             \
    do {                                                                        \
        if (!(cond)) {                                                          \
            std::printf("CHECK FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                         \
        }                                                                       \
    } while (0)

template <typename E>
static bool throws(const std::function<void()>& f, const std::string& needle = "") {
    try {
        f();
    } catch (const E& e) {
        return needle.empty() || std::string(e.what()).find(needle) != std::string::npos;
    } catch (...) {
        return false;
    }
    return false;
}

static int validate(const std::string& dir) {
    const mesh::MeshData sphere = mesh::SphericalSurface(1.0f, 1).getMeshData();
    CHECK(sphere.faces.size() == 80 && sphere.edges.size() == 120 && sphere.vertices.size() == 42);
    // edges rebuilt from the faces keep the generator's first-encounter numbering
    mesh::MeshData copy = sphere;
    mesh::build_edges(copy);
    bool same = copy.faces_edges.size() == sphere.faces_edges.size();
    for (std::size_t f = 0; same && f < copy.faces_edges.size(); ++f)
        same = copy.faces_edges[f].x == sphere.faces_edges[f].x && copy.faces_edges[f].y == sphere.faces_edges[f].y &&
               copy.faces_edges[f].z == sphere.faces_edges[f].z;
    CHECK(same);
    // OBJ round trip and the closed-surface checks
    const std::string obj = dir + "/sphere.obj";
    {
        std::ofstream o(obj);
        o << "# icosphere level 1\n";
        for (const auto& v : sphere.vertices) o << "v " << v.x << " " << v.y << " " << v.z << "\n";
        for (const auto& f : sphere.faces) o << "f " << f.x + 1 << "/1 " << f.y + 1 << " " << -(int)sphere.vertices.size() + f.z << "\n";
    }
    const mesh::MeshData m = mesh::load_obj(obj);
    CHECK(m.faces.size() == 80 && m.edges.size() == 120);
    CHECK(m.faces[5].x == sphere.faces[5].x && m.faces[5].z == sphere.faces[5].z);
    const std::string bad = dir + "/open.obj";
    {
        std::ofstream o(bad);
        o << "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n";
    }
    CHECK(throws<std::runtime_error>([&] { mesh::load_obj(bad); }, "boundary edge"));
    const std::string quad = dir + "/quad.obj";
    {
        std::ofstream o(quad);
        o << "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf 1 2 4 3\n";
    }
    CHECK(throws<std::runtime_error>([&] { mesh::load_obj(quad); }, "only triangular faces"));
    CHECK(throws<std::runtime_error>([&] { mesh::load_obj(dir + "/missing.obj"); }, "Failed to open"));
    // tensors.bin: complex128 and complex64 forms
    std::vector<std::complex<real_t>> t(80 * 32);
    for (std::size_t i = 0; i < t.size(); ++i) t[i] = std::complex<real_t>(0.25 * i, -0.5 * i);
    io::write_tensors(dir + "/t64.bin", t);
    CHECK(io::read_tensors(dir + "/t64.bin", 80) == t);
    {
        std::ofstream o(dir + "/t32.bin", std::ios::binary);
        for (const auto& z : t) {
            const float a[2] = {(float)z.real(), (float)z.imag()};
            o.write(reinterpret_cast<const char*>(a), sizeof a);
        }
    }
    CHECK(io::read_tensors(dir + "/t32.bin", 80) == t);
    CHECK(throws<std::runtime_error>([&] { io::read_tensors(dir + "/t64.bin", 79); }, "size mismatch"));
    io::write_currents(dir + "/c.bin", t);
    CHECK(io::read_currents(dir + "/c.bin") == t);
    // option and source validation happen before any device work
    solvers::GstcSieOptions o;
    const double f = 1e8;
    CHECK(throws<std::invalid_argument>([&] { excitation::prepare_source(excitation::PlaneWaveExcitation{}, 0.0, {}, {}); },
                                        "source frequency"));
    excitation::PlaneWaveExcitation pw;
    pw.electric_amplitude = make_vec3(1, 0, 1);
    CHECK(throws<std::invalid_argument>([&] { excitation::prepare_source(pw, f, {}, {}); }, "transverse"));
    CHECK(throws<std::invalid_argument>([&] { excitation::prepare_source(excitation::DipoleExcitation{}, f, {}, {}); }, "dipole"));
    const auto src = excitation::prepare_source(excitation::PlaneWaveExcitation{}, f, {}, {});
    CHECK(std::abs(src.plane_wave().k - 2 * M_PI * f / 299792458.0) < 1e-6 * src.plane_wave().k);
    o.frequency_hz = f;
    CHECK(throws<std::invalid_argument>([&] { solvers::solve_gstc_dense(sphere, std::vector<std::complex<real_t>>(31), o, src); },
                                        "GSTC tensor count mismatch: expected 2560, got 31"));
    o.quadrature_rule = 21;
    CHECK(throws<std::invalid_argument>([&] { solvers::solve_gstc_dense(sphere, t, o, src); }, "Dunavant quadrature rule"));
    o.quadrature_rule = 5;
    solvers::GstcSieOptions o2 = o;
    o2.frequency_hz = 2 * f;
    CHECK(throws<std::invalid_argument>([&] { solvers::solve_gstc_dense(sphere, t, o2, src); }, "does not match"));
    CHECK(throws<std::invalid_argument>([&] { solvers::solve_gstc_dense_batch(sphere, t, o, {}); }, "at least one excitation"));
    solvers::GstcExcitationCase wrong;
    wrong.interior = src;  // an exterior source in the interior slot
    CHECK(throws<std::invalid_argument>([&] { solvers::solve_gstc_dense_batch(sphere, t, o, {wrong}); },
                                        "SIE batch case 0: excitation does not match its regional slot"));
    solvers::PecOptions po;
    po.frequency_hz = f;
    po.formulation = solvers::PecFormulation::NMfie;
    const auto src18 = excitation::prepare_source(excitation::PlaneWaveExcitation{}, f, {}, {}, kCodata2018Eps0, kCodata2018Mu0);
    CHECK(throws<std::invalid_argument>([&] { solvers::solve_pec_dense(sphere, po, src18); }, "T-EFIE"));
    solvers::DielectricOptions dopt;
    dopt.frequency_hz = f;
    dopt.formulation = solvers::DielectricFormulation::N_Muller;
    CHECK(throws<std::invalid_argument>([&] { solvers::solve_dielectric_dense(sphere, dopt, src18); }, "T-PMCHWT"));
    CHECK(throws<std::invalid_argument>([&] { gpu::DenseLuSolver lu(0); }, "dimension must be positive"));
        // field evaluation: argument checks (host side, before any device work)
    {
        using solvers::evaluate_surface_integral_fields;
        const std::vector<std::complex<real_t>> cur(sphere.edges.size(), std::complex<real_t>(1, 0));
        const std::vector<vec3_t> pts = {make_vec3(0, 0, 2)};
        CHECK(throws<std::invalid_argument>([&] { evaluate_surface_integral_fields(sphere, cur, cur, pts, 0.0, {}); },
                                            "field-evaluation frequency must be positive and finite"));
        IsotropicMedium bad;
        bad.relative_permittivity = -1;
        CHECK(throws<std::invalid_argument>([&] { evaluate_surface_integral_fields(sphere, cur, cur, pts, 1e9, bad); },
                                            "relative permittivity and permeability must be positive"));
        CHECK(throws<std::invalid_argument>(
            [&] { evaluate_surface_integral_fields(sphere, cur, cur, pts, 1e9, {}, 21); },
            "Dunavant quadrature rule must be in [1, 20]"));
        const std::vector<vec3_t> nanpt = {make_vec3(0, 0, 2), make_vec3(std::nan(""), 0, 0)};
        CHECK(throws<std::invalid_argument>([&] { evaluate_surface_integral_fields(sphere, cur, cur, nanpt, 1e9, {}); },
                                            "non-finite observation point at index 1"));
        CHECK(evaluate_surface_integral_fields(sphere, cur, cur, {}, 1e9, {}).empty());
        const std::vector<std::complex<real_t>> shortc(5);
        CHECK(throws<std::invalid_argument>([&] { evaluate_surface_integral_fields(sphere, shortc, cur, pts, 1e9, {}); },
                                            "one electric and magnetic coefficient per mesh edge"));
    }
    std::printf("validate: %d failure(s)\n", failures);
    return failures ? 1 : 0;
}

static void dump(const std::string& path, const std::vector<std::complex<real_t>>& v) { io::write_currents(path, v); }

static int solve(const std::string& dir, double f, int rule) {
    const mesh::MeshData sphere = mesh::SphericalSurface(1.0f, 1).getMeshData();
    const std::vector<std::complex<real_t>> tensors = io::read_tensors(dir + "/tensors.bin", sphere.faces.size());
    solvers::GstcSieOptions o;
    o.frequency_hz = f;
    o.quadrature_rule = rule;
    const auto ext = excitation::prepare_source(excitation::PlaneWaveExcitation{}, f, o.interior, o.exterior);
    excitation::PlaneWaveExcitation pin;
    pin.region = excitation::SourceRegion::interior;
    pin.propagation_direction = make_vec3(0, 1, 0);
    pin.electric_amplitude = make_vec3(0, 0, 2);
    const auto inner = excitation::prepare_source(pin, f, o.interior, o.exterior);
    auto r64 = solvers::solve_gstc_dense(sphere, tensors, o, ext);
    dump(dir + "/gstc_fp64.bin", r64.physical_currents);
    CHECK(r64.metrics.residual_computed && r64.metrics.relative_residual < 1e-12);
    CHECK(r64.metrics.matrix_dimension == 480 && r64.metrics.maximum_j2_coefficient > 0);
    solvers::GstcSieOptions om = o;
    om.precision = solvers::SolvePrecision::mixed;
    auto rmx = solvers::solve_gstc_dense(sphere, tensors, om, ext);
    dump(dir + "/gstc_mixed.bin", rmx.physical_currents);
    CHECK(rmx.metrics.relative_residual < 1e-11);
    solvers::GstcExcitationCase c0, c1;
    c0.exterior = ext;
    c1.interior = inner;
    c1.exterior = ext;
    auto batch = solvers::solve_gstc_dense_batch(sphere, tensors, o, {c0, c1});
    CHECK(batch.cases.size() == 2 && batch.metrics.case_count == 2);
    dump(dir + "/gstc_batch0.bin", batch.cases[0].physical_currents);
    dump(dir + "/gstc_batch1.bin", batch.cases[1].physical_currents);
    CHECK(batch.cases[1].metrics.relative_residual < 1e-12);
    // PEC T-EFIE and T-PMCHWT with the CODATA-2018 defaults of their option structs
    solvers::PecOptions po;
    po.frequency_hz = f;
    po.quadrature_rule = rule;
    const auto src18 = excitation::prepare_source(excitation::PlaneWaveExcitation{}, f, {}, po.exterior,
                                                  po.vacuum_permittivity, po.vacuum_permeability);
    po.capture_assembled_matrix = true;
    auto pec = solvers::solve_pec_dense(sphere, po, src18);
    dump(dir + "/pec.bin", pec.exterior_current);
    {
        // fields of the PEC currents at far, near-surface and interior points
        const std::vector<vec3_t> pts = {make_vec3(0, 0, 3), make_vec3(1.5, -0.5, 0.25), make_vec3(0.2, 0.1, -0.3),
                                         make_vec3(0, 0, 0.6), make_vec3(0.5, 0.5, 0.55)};
        const std::vector<std::complex<real_t>> zero(pec.exterior_current.size());
        const auto smp = solvers::evaluate_surface_integral_fields(sphere, pec.exterior_current, zero, pts, f, {},
                                                                   rule, po.vacuum_permittivity,
                                                                   po.vacuum_permeability);
        std::vector<std::complex<real_t>> eh;
        for (const auto& q : smp)
            for (int c = 0; c < 3; ++c) eh.push_back(q.electric[c]);
        for (const auto& q : smp)
            for (int c = 0; c < 3; ++c) eh.push_back(q.magnetic[c]);
        dump(dir + "/pec_fields.bin", eh);
    }
    CHECK(pec.assembled_matrix.size() == 120 * 120 && pec.metrics.relative_residual < 1e-12);
    solvers::DielectricOptions dopt;
    dopt.frequency_hz = f;
    dopt.quadrature_rule = rule;
    dopt.interior.relative_permittivity = 4.0;
    const auto srcd = excitation::prepare_source(excitation::PlaneWaveExcitation{}, f, dopt.interior, dopt.exterior,
                                                 dopt.vacuum_permittivity, dopt.vacuum_permeability);
    auto die = solvers::solve_dielectric_dense(sphere, dopt, srcd);
    dump(dir + "/pmchwt.bin", die.physical_currents);
    CHECK(die.metrics.relative_residual < 1e-12 && die.metrics.maximum_m_coefficient > 0);
    {
        // PMCHWT batch: FP64 cases match the single solves; six waves in mixed
        // precision (blocked refinement) agree with FP64
        std::vector<excitation::PreparedExcitation> srcs;
        for (int wv = 0; wv < 6; ++wv) {
            excitation::PlaneWaveExcitation pw;
            const double th = 0.4 * wv;
            pw.propagation_direction = make_vec3(std::sin(th), 0, std::cos(th));
            pw.electric_amplitude = (wv % 2) ? make_vec3(0, 1, 0) : make_vec3(std::cos(th), 0, -std::sin(th));
            srcs.push_back(excitation::prepare_source(pw, f, dopt.interior, dopt.exterior, dopt.vacuum_permittivity,
                                                      dopt.vacuum_permeability));
        }
        srcs[0] = srcd;
        auto b64 = solvers::solve_dielectric_dense_batch(sphere, dopt, srcs);
        CHECK(b64.cases.size() == 6);
        {
            // six columns take the blocked multi-RHS solve: equal to the single solve to rounding
            double num = 0, den = 0;
            for (std::size_t i = 0; i < die.physical_currents.size(); ++i) {
                num += std::norm(b64.cases[0].physical_currents[i] - die.physical_currents[i]);
                den += std::norm(die.physical_currents[i]);
            }
            CHECK(std::sqrt(num / den) < 1e-12);
        }
        solvers::DielectricOptions dm = dopt;
        dm.precision = solvers::SolvePrecision::mixed;
        auto bmx = solvers::solve_dielectric_dense_batch(sphere, dm, srcs);
        double worst = 0;
        for (int c = 0; c < 6; ++c) {
            double num = 0, den = 0;
            for (std::size_t i = 0; i < b64.cases[c].physical_currents.size(); ++i) {
                num += std::norm(bmx.cases[c].physical_currents[i] - b64.cases[c].physical_currents[i]);
                den += std::norm(b64.cases[c].physical_currents[i]);
            }
            worst = std::max(worst, std::sqrt(num / den));
            CHECK(bmx.cases[c].metrics.relative_residual < 1e-11);
        }
        CHECK(worst < 1e-10);
        CHECK(throws<std::invalid_argument>([&] { solvers::solve_dielectric_dense_batch(sphere, dopt, {}); },
                                            "at least one excitation"));
    }
    // operator level: GMRES on the stored PEC matrix against the LU solution
    {
        const int n = 120;
        std::complex<real_t> *Z, *b, *x;
        if (hipMalloc((void**)&Z, sizeof(std::complex<real_t>) * n * n) != 0 || hipMalloc((void**)&b, sizeof(std::complex<real_t>) * n) != 0 ||
            hipMalloc((void**)&x, sizeof(std::complex<real_t>) * n) != 0)
            throw std::runtime_error("allocation");
        std::vector<std::complex<real_t>> hb(n);
        for (int i = 0; i < n; ++i) hb[i] = std::complex<real_t>(std::cos(0.3 * i), std::sin(0.7 * i));
        (void)hipMemcpy(Z, pec.assembled_matrix.data(), sizeof(std::complex<real_t>) * n * n, hipMemcpyHostToDevice);
        (void)hipMemcpy(b, hb.data(), sizeof(std::complex<real_t>) * n, hipMemcpyHostToDevice);
        (void)hipMemset(x, 0, sizeof(std::complex<real_t>) * n);
        gpu::GMRESSolver<gpu::DenseMatVec> gm(n, gpu::DenseMatVec{Z, n, n});
        double res = 0;
        int its = 0;
        std::vector<double> hist;
        const int info = gm.solve(b, x, 60, 5, 1e-12, &res, &its, &hist);
        CHECK(info == 0 && its > 0 && (int)hist.size() == its);
        std::vector<std::complex<real_t>> xg(n), xl = hb;
        (void)hipMemcpy(xg.data(), x, sizeof(std::complex<real_t>) * n, hipMemcpyDeviceToHost);
        (void)hipMemcpy(b, hb.data(), sizeof(std::complex<real_t>) * n, hipMemcpyHostToDevice);
        gpu::solve_dense_lu(n, Z, b);
        (void)hipMemcpy(xl.data(), b, sizeof(std::complex<real_t>) * n, hipMemcpyDeviceToHost);
        double num = 0, den = 0;
        for (int i = 0; i < n; ++i) {
            num += std::norm(xg[i] - xl[i]);
            den += std::norm(xl[i]);
        }
        CHECK(std::sqrt(num / den) < 1e-9);
        (void)hipFree(Z);
        (void)hipFree(b);
        (void)hipFree(x);
    }
    std::printf("solve: %d failure(s)\n", failures);
    return failures ? 1 : 0;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s validate <tmpdir> | solve <dir> <frequency_hz> <rule>\n", argv[0]);
        return 2;
    }
    try {
        const std::string mode = argv[1];
        if (mode == "validate") return validate(argv[2]);
        if (mode == "solve" && argc >= 5) return solve(argv[2], std::strtod(argv[3], nullptr), std::atoi(argv[4]));
    } catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 1;
    }
    return 2;
}
