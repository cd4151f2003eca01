// solve_gstc_dense / solve_gstc_dense_batch / solve_pec_dense /
// solve_dielectric_dense over the C ABI, with the reference drivers' contract
// (src/sie_mom/solvers/gstc_sie.cu:71-720, pec_solver.cu:20-488,
// dielectric/dielectric_solver.cu:100-740): validation order and messages,
// metrics, residual of the re-assembled operator, physical-current recovery.
#include <algorithm>
#include <limits>

#include "host_common.hpp"
#include "sie_mom/gpu/dense_lu_solver.hpp"
#include "sie_mom/solvers/dielectric/dielectric_solver.hpp"
#include "sie_mom/solvers/gstc_sie.hpp"
#include "sie_mom/solvers/pec_solver.hpp"

namespace sie_mom {
namespace solvers {
namespace {

using host::Clock;
using host::DeviceArray;
using host::seconds_since;

// Kernel launch parameters cast to float as the reference casts them
// (gstc_sie.cu:281-290, pec_solver.cu:352-359, dielectric_solver.cu:335-347).
sie_calderon_params make_params(int rule, double k1, double k2, const IsotropicMedium& interior,
                                const IsotropicMedium& exterior, double mu0, double eps0, double omega, double thr) {
    sie_calderon_params p;
    p.rule = rule;
    p.k1 = (float)k1;
    p.k2 = (float)k2;
    p.mu_r1 = (float)interior.relative_permeability;
    p.mu_r2 = (float)exterior.relative_permeability;
    p.eps_r1 = (float)interior.relative_permittivity;
    p.eps_r2 = (float)exterior.relative_permittivity;
    p.mu0 = (float)mu0;
    p.eps0 = (float)eps0;
    p.omega = (float)omega;
    p.threshold = (float)thr;
    return p;
}

double lambda0_of(double omega, double mu0, double eps0) { return 2.0 * std::acos(-1.0) / (omega * std::sqrt(mu0 * eps0)); }

void validate_common(const mesh::MeshData& mesh, double frequency, const IsotropicMedium& interior,
                     const IsotropicMedium& exterior, double eps0, double mu0, int rule, double thr) {
    if (mesh.vertices.empty() || mesh.faces.empty() || mesh.edges.empty())
        throw std::invalid_argument("SIE mesh must contain vertices, faces, and edges");
    if (mesh.faces_edges.size() != mesh.faces.size())
        throw std::invalid_argument("mesh face-to-edge mapping size does not match face count");
    if (!(frequency > 0)) throw std::invalid_argument("SIE frequency must be positive");
    if (!(interior.relative_permittivity > 0) || !(interior.relative_permeability > 0) ||
        !(exterior.relative_permittivity > 0) || !(exterior.relative_permeability > 0))
        throw std::invalid_argument("relative permittivities and permeabilities must be positive");
    if (!(eps0 > 0) || !(mu0 > 0)) throw std::invalid_argument("vacuum permittivity and permeability must be positive");
    if (rule < 1 || rule > 20) throw std::invalid_argument("Dunavant quadrature rule must be in [1, 20]");
    if (!(thr > 0)) throw std::invalid_argument("singularity threshold in lambda0 must be positive");
}

void check_dimension(std::size_t dim) {
    if (dim > (std::size_t)std::numeric_limits<int>::max())
        throw std::overflow_error("SIE matrix dimension exceeds the solver int limit");
    if (dim > std::numeric_limits<std::size_t>::max() / dim) throw std::overflow_error("dense SIE matrix element count overflows size_t");
    if (dim * dim > std::numeric_limits<std::size_t>::max() / sizeof(std::complex<real_t>))
        throw std::overflow_error("dense SIE matrix byte count overflows size_t");
}

void validate_gstc(const mesh::MeshData& mesh, const std::vector<std::complex<real_t>>& tensors, const GstcSieOptions& o) {
    validate_common(mesh, o.frequency_hz, o.interior, o.exterior, o.vacuum_permittivity, o.vacuum_permeability,
                    o.quadrature_rule, o.singularity_threshold_lambda0);
    if (o.basis != GstcBasisType::face && o.basis != GstcBasisType::spherical)
        throw std::invalid_argument("unsupported GSTC basis type");
    const std::size_t records = o.tensors_are_uniform ? 1 : mesh.faces.size();
    if (records > std::numeric_limits<std::size_t>::max() / 32) throw std::overflow_error("GSTC tensor count overflows size_t");
    if (tensors.size() != records * 32)
        throw std::invalid_argument("GSTC tensor count mismatch: expected " + std::to_string(records * 32) + ", got " +
                                    std::to_string(tensors.size()));
    if (mesh.edges.size() > std::numeric_limits<std::size_t>::max() / 4) throw std::overflow_error("SIE matrix dimension overflows size_t");
    check_dimension(4 * mesh.edges.size());
}

// Per-block maxima of |coefficient| and the finiteness check of the reference
// (validate_and_summarize_physical_currents, gstc_sie.cu:474-503).
std::vector<double> block_maxima(const std::vector<std::complex<real_t>>& c, std::size_t ne, int blocks, const char* what) {
    std::vector<double> mx(blocks, 0.0);
    for (int b = 0; b < blocks; ++b)
        for (std::size_t e = 0; e < ne; ++e) {
            const std::size_t i = b * ne + e;
            if (!std::isfinite(c[i].real()) || !std::isfinite(c[i].imag()))
                throw std::runtime_error(std::string(what) + std::to_string(i));
            mx[b] = std::max(mx[b], std::hypot(c[i].real(), c[i].imag()));
        }
    return mx;
}

// r = Z x - b on the device, |r| on the host (long double accumulation).
void residual(int n, const DeviceArray<std::complex<real_t>>& Z, const std::complex<real_t>* x_dev, const std::vector<std::complex<real_t>>& b,
              double rhs_norm, double& abs_res, double& rel_res) {
    DeviceArray<std::complex<real_t>> r(n);
    r.upload(b.data(), n);
    detail::check(sie_zgemv(0, n, n, 1.0, reinterpret_cast<const double*>(Z.data()), n,
                            reinterpret_cast<const double*>(x_dev), -1.0, reinterpret_cast<double*>(r.data()), nullptr));
    std::vector<std::complex<real_t>> h(n);
    r.download(h.data(), n);
    abs_res = host::host_norm(h);
    rel_res = rhs_norm > 0 ? abs_res / rhs_norm : std::numeric_limits<double>::infinity();
}

struct GstcDevice {
    host::MeshHandle mesh;
    DeviceArray<std::complex<real_t>> tensors;
    sie_calderon_params params;
    GstcDevice(const mesh::MeshData& m, const std::vector<std::complex<real_t>>& t, const sie_calderon_params& p)
        : mesh(m), tensors(t.size()), params(p) {
        tensors.upload(t.data(), t.size());
    }
    void assemble(const GstcSieOptions& o, std::complex<real_t>* Z, int n) const {
        detail::check(sie_assemble_gstc(mesh.get(), &params, reinterpret_cast<const double*>(tensors.data()),
                                        o.tensors_are_uniform ? 1 : 0, (int)o.basis, SIE_ASM_CALDERON | SIE_ASM_MASS,
                                        reinterpret_cast<double*>(Z), n, nullptr));
    }
    // Tested incident fields of the given waves, summed into one column.
    void rhs(const GstcSieOptions& o, const std::vector<sie_plane_wave>& waves, std::complex<real_t>* col, int n) const {
        DeviceArray<std::complex<real_t>> tmp((size_t)n * waves.size());
        detail::check(sie_rhs_plane_waves(mesh.get(), SIE_RHS_GSTC, o.quadrature_rule, waves.data(), (int)waves.size(),
                                          (float)o.vacuum_permeability, (float)o.vacuum_permittivity,
                                          o.vacuum_permeability, o.vacuum_permittivity,
                                          reinterpret_cast<double*>(waves.size() == 1 ? col : tmp.data()), n, nullptr));
        if (waves.size() > 1) {
            std::vector<std::complex<real_t>> ones(waves.size(), std::complex<real_t>(1.0));
            DeviceArray<std::complex<real_t>> w(ones.size());
            w.upload(ones.data(), ones.size());
            detail::check(sie_zgemv(0, n, (int)waves.size(), 1.0, reinterpret_cast<const double*>(tmp.data()), n,
                                    reinterpret_cast<const double*>(w.data()), 0.0, reinterpret_cast<double*>(col),
                                    nullptr));
        }
        host::hip_ok(hipDeviceSynchronize(), "GSTC right-hand side");
    }
};

sie_calderon_params gstc_params(const GstcSieOptions& o, const PhysicalParameters& ph) {
    return make_params(o.quadrature_rule, ph.k_interior, ph.k_exterior, o.interior, o.exterior, o.vacuum_permeability,
                       o.vacuum_permittivity, ph.omega, ph.singularity_threshold);
}

void recover(std::vector<std::complex<real_t>>& c, std::size_t ne, double eta0) {
    for (std::size_t e = 0; e < ne; ++e) {
        c[ne + e] *= eta0;
        c[3 * ne + e] *= eta0;
    }
}

// Factorisation and solves of the assembled system; the solutions replace
// the right-hand-side columns of X.
class ColumnSolver {
public:
    ColumnSolver(const GstcSieOptions& o, DeviceArray<std::complex<real_t>>& Z, int n) : o_(o), Z_(Z), n_(n) {}
    void factorize() {
        if (o_.precision == SolvePrecision::fp64) {
            lu_.allocate(n_);
            lu_.factorize(Z_.data());
        } else {
            mixed_.factorize(Z_.data(), n_);
        }
        host::hip_ok(hipDeviceSynchronize(), "LU factorisation");
    }
    void solve(DeviceArray<std::complex<real_t>>& X, int nrhs) {
        if (o_.precision == SolvePrecision::fp64) {
            lu_.solve(gpu::Op::N, nrhs, Z_.data(), X.data(), n_);
            return;
        }
        DeviceArray<std::complex<real_t>> B((size_t)n_ * nrhs);
        host::hip_ok(hipMemcpy(B.data(), X.data(), sizeof(std::complex<real_t>) * n_ * nrhs, hipMemcpyDeviceToDevice), "copy RHS");
        const auto be = mixed_.solve(gpu::Op::N, nrhs, B.data(), n_, X.data(), n_, o_.refinement_tolerance,
                                     o_.refinement_max_sweeps);
        for (double r : be)
            if (!(r <= o_.refinement_tolerance))
                throw std::runtime_error("iterative refinement did not reach the requested backward error (" +
                                         std::to_string(r) + ")");
    }

private:
    const GstcSieOptions& o_;
    DeviceArray<std::complex<real_t>>& Z_;
    int n_;
    gpu::DenseLuSolver lu_;
    gpu::MixedLuSolver mixed_;
};

}  // namespace

PhysicalParameters make_physical_parameters(const GstcSieOptions& o) {
    const double pi = std::acos(-1.0);
    const double omega = 2.0 * pi * o.frequency_hz;
    const double ei = o.interior.relative_permittivity * o.vacuum_permittivity;
    const double mi = o.interior.relative_permeability * o.vacuum_permeability;
    const double ee = o.exterior.relative_permittivity * o.vacuum_permittivity;
    const double me = o.exterior.relative_permeability * o.vacuum_permeability;
    const double eta0 = std::sqrt(o.vacuum_permeability / o.vacuum_permittivity);
    const double c0 = 1.0 / std::sqrt(o.vacuum_permeability * o.vacuum_permittivity);
    return {omega, omega * std::sqrt(mi * ei), omega * std::sqrt(me * ee), eta0,
            o.singularity_threshold_lambda0 * c0 / o.frequency_hz};
}

GstcSieResult solve_gstc_dense(const mesh::MeshData& mesh, const std::vector<std::complex<real_t>>& gstc_tensors,
                               const GstcSieOptions& options, const excitation::PreparedExcitation& source) {
    validate_gstc(mesh, gstc_tensors, options);
    source.validate_environment(options.frequency_hz, options.interior, options.exterior, options.vacuum_permittivity,
                                options.vacuum_permeability);
    host::require_device();
    GstcSieResult result;
    auto& mt = result.metrics;
    mt.vertex_count = mesh.vertices.size();
    mt.face_count = mesh.faces.size();
    mt.edge_count = mesh.edges.size();
    mt.matrix_dimension = 4 * mesh.edges.size();
    mt.dense_matrix_bytes = mt.matrix_dimension * mt.matrix_dimension * sizeof(std::complex<real_t>);
    const int n = (int)mt.matrix_dimension;
    const auto ph = make_physical_parameters(options);
    auto t0 = Clock::now();
    GstcDevice dev(mesh, gstc_tensors, gstc_params(options, ph));
    DeviceArray<std::complex<real_t>> Z((size_t)n * n), x(n);
    mt.device_preparation_seconds = seconds_since(t0);
    t0 = Clock::now();
    dev.assemble(options, Z.data(), n);
    dev.rhs(options, {source.plane_wave()}, x.data(), n);
    mt.assembly_seconds = seconds_since(t0);
    std::vector<std::complex<real_t>> b(n);
    x.download(b.data(), n);
    mt.rhs_norm = host::host_norm(b);
    if (!(mt.rhs_norm > 0.0)) throw std::runtime_error("assembled incident right-hand side has zero norm");
    t0 = Clock::now();
    ColumnSolver cs(options, Z, n);
    cs.factorize();
    cs.solve(x, 1);
    mt.solve_seconds = seconds_since(t0);
    std::vector<std::complex<real_t>> cur(n);
    x.download(cur.data(), n);
    if (options.compute_residual) {
        t0 = Clock::now();
        if (options.precision == SolvePrecision::fp64) dev.assemble(options, Z.data(), n);  // LU replaced Z
        residual(n, Z, x.data(), b, mt.rhs_norm, mt.absolute_residual, mt.relative_residual);
        mt.residual_computed = true;
        mt.residual_seconds = seconds_since(t0);
    }
    recover(cur, mesh.edges.size(), ph.eta0);
    const auto mx = block_maxima(cur, mesh.edges.size(), 4, "non-finite physical current coefficient at index ");
    mt.maximum_j1_coefficient = mx[0];
    mt.maximum_m1_coefficient = mx[1];
    mt.maximum_j2_coefficient = mx[2];
    mt.maximum_m2_coefficient = mx[3];
    result.physical_currents = std::move(cur);
    return result;
}

GstcSieBatchResult solve_gstc_dense_batch(const mesh::MeshData& mesh, const std::vector<std::complex<real_t>>& gstc_tensors,
                                          const GstcSieOptions& options,
                                          const std::vector<GstcExcitationCase>& cases) {
    validate_gstc(mesh, gstc_tensors, options);
    const std::size_t dim = 4 * mesh.edges.size();
    if (cases.empty()) throw std::invalid_argument("SIE batch must contain at least one excitation case");
    if (cases.size() > (std::size_t)std::numeric_limits<int>::max())
        throw std::overflow_error("SIE batch case count exceeds the solver int limit");
    if (cases.size() > std::numeric_limits<std::size_t>::max() / dim / sizeof(std::complex<real_t>))
        throw std::overflow_error("SIE batch RHS byte count overflows size_t");
    std::vector<std::vector<sie_plane_wave>> waves(cases.size());
    for (std::size_t i = 0; i < cases.size(); ++i) {
        const std::string prefix = "SIE batch case " + std::to_string(i) + ": ";
        const auto& c = cases[i];
        if (!c.interior && !c.exterior) throw std::invalid_argument(prefix + "at least one incident region is required");
        const std::pair<const std::optional<excitation::PreparedExcitation>*, excitation::SourceRegion> slots[2] = {
            {&c.interior, excitation::SourceRegion::interior}, {&c.exterior, excitation::SourceRegion::exterior}};
        for (const auto& slot : slots) {
            if (!*slot.first) continue;
            const auto& src = **slot.first;
            if (src.region() != slot.second) throw std::invalid_argument(prefix + "excitation does not match its regional slot");
            try {
                src.validate_environment(options.frequency_hz, options.interior, options.exterior,
                                         options.vacuum_permittivity, options.vacuum_permeability);
            } catch (const std::invalid_argument& e) {
                throw std::invalid_argument(prefix + e.what());
            }
            waves[i].push_back(src.plane_wave());
        }
    }
    host::require_device();
    GstcSieBatchResult result;
    result.cases.resize(cases.size());
    auto& mt = result.metrics;
    mt.vertex_count = mesh.vertices.size();
    mt.face_count = mesh.faces.size();
    mt.edge_count = mesh.edges.size();
    mt.matrix_dimension = dim;
    mt.dense_matrix_bytes = dim * dim * sizeof(std::complex<real_t>);
    mt.case_count = cases.size();
    const int n = (int)dim, nc = (int)cases.size();
    const auto ph = make_physical_parameters(options);
    auto t0 = Clock::now();
    GstcDevice dev(mesh, gstc_tensors, gstc_params(options, ph));
    DeviceArray<std::complex<real_t>> Z((size_t)n * n), X((size_t)n * nc);
    mt.device_preparation_seconds = seconds_since(t0);
    t0 = Clock::now();
    dev.assemble(options, Z.data(), n);
    host::hip_ok(hipDeviceSynchronize(), "GSTC assembly");
    mt.matrix_assembly_seconds = seconds_since(t0);
    std::vector<std::vector<std::complex<real_t>>> rhs(cases.size(), std::vector<std::complex<real_t>>(n));
    for (int i = 0; i < nc; ++i) {
        auto& cm = result.cases[i].metrics;
        const auto tr = Clock::now();
        dev.rhs(options, waves[i], X.data() + (size_t)i * n, n);
        cm.rhs_assembly_seconds = seconds_since(tr);
        X.download(rhs[i].data(), n, (size_t)i * n);
        cm.rhs_norm = host::host_norm(rhs[i]);
        if (!(cm.rhs_norm > 0.0) || !std::isfinite(cm.rhs_norm))
            throw std::runtime_error("SIE batch case " + std::to_string(i) + ": incident RHS norm must be finite and nonzero");
    }
    t0 = Clock::now();
    ColumnSolver cs(options, Z, n);
    cs.factorize();
    mt.factorization_seconds = seconds_since(t0);
    t0 = Clock::now();
    cs.solve(X, nc);
    mt.solve_seconds = seconds_since(t0);
    if (options.compute_residual && options.precision == SolvePrecision::fp64) {
        t0 = Clock::now();
        dev.assemble(options, Z.data(), n);
        host::hip_ok(hipDeviceSynchronize(), "GSTC re-assembly");
        mt.residual_matrix_assembly_seconds = seconds_since(t0);
    }
    for (int i = 0; i < nc; ++i) {
        auto& cr = result.cases[i];
        std::vector<std::complex<real_t>> cur(n);
        X.download(cur.data(), n, (size_t)i * n);
        if (options.compute_residual) {
            const auto tr = Clock::now();
            residual(n, Z, X.data() + (size_t)i * n, rhs[i], cr.metrics.rhs_norm, cr.metrics.absolute_residual,
                     cr.metrics.relative_residual);
            cr.metrics.residual_computed = true;
            cr.metrics.residual_seconds = seconds_since(tr);
        }
        recover(cur, mesh.edges.size(), ph.eta0);
        const auto mx = block_maxima(cur, mesh.edges.size(), 4, "non-finite physical current coefficient at index ");
        cr.metrics.maximum_j1_coefficient = mx[0];
        cr.metrics.maximum_m1_coefficient = mx[1];
        cr.metrics.maximum_j2_coefficient = mx[2];
        cr.metrics.maximum_m2_coefficient = mx[3];
        cr.physical_currents = std::move(cur);
    }
    return result;
}

PecResult solve_pec_dense(const mesh::MeshData& mesh, const PecOptions& o, const excitation::PreparedExcitation& source) {
    if (mesh.vertices.empty() || mesh.faces.empty() || mesh.edges.empty() || mesh.faces_edges.size() != mesh.faces.size())
        throw std::invalid_argument("PEC mesh needs vertices, faces, edges, and a face-edge map");
    if (!(o.frequency_hz > 0) || !(o.exterior.relative_permittivity > 0) || !(o.exterior.relative_permeability > 0) ||
        !(o.vacuum_permittivity > 0) || !(o.vacuum_permeability > 0))
        throw std::invalid_argument("PEC frequency and medium constants must be positive");
    if (o.quadrature_rule < 1 || o.quadrature_rule > 20) throw std::invalid_argument("Dunavant rule must be in [1, 20]");
    if (o.line_quadrature_points < 1 || o.line_quadrature_points > 20)
        throw std::invalid_argument("Gauss-Legendre point count must be in [1, 20]");
    if (o.formulation != PecFormulation::NEfie && o.ntl_form == PecNtlForm::Line)
        throw std::invalid_argument("line nTL form requires N-EFIE");
    if (!(o.singularity_threshold_lambda0 > 0)) throw std::invalid_argument("singularity threshold in lambda0 must be positive");
    if (source.region() != excitation::SourceRegion::exterior)
        throw std::invalid_argument("PEC solver requires an exterior (region 2) source");
    if (o.formulation != PecFormulation::TEfie)
        throw std::invalid_argument("only the T-EFIE PEC formulation is provided by this library");
    if (mesh.edges.size() > (std::size_t)std::numeric_limits<int>::max())
        throw std::invalid_argument("PEC edge count exceeds the dense solver integer limit");
    source.validate_environment(o.frequency_hz, o.exterior, o.exterior, o.vacuum_permittivity, o.vacuum_permeability);
    host::require_device();
    check_dimension(mesh.edges.size());
    PecResult result;
    auto& mt = result.metrics;
    mt.vertex_count = mesh.vertices.size();
    mt.face_count = mesh.faces.size();
    mt.edge_count = mesh.edges.size();
    mt.matrix_dimension = mesh.edges.size();
    mt.dense_matrix_bytes = mt.matrix_dimension * mt.matrix_dimension * sizeof(std::complex<real_t>);
    const int n = (int)mt.matrix_dimension;
    const double omega = 2.0 * std::acos(-1.0) * o.frequency_hz;
    const double ke = omega * std::sqrt(o.exterior.relative_permeability * o.vacuum_permeability *
                                        o.exterior.relative_permittivity * o.vacuum_permittivity);
    const double thr = o.singularity_threshold_lambda0 * lambda0_of(omega, o.vacuum_permeability, o.vacuum_permittivity);
    const sie_calderon_params p = make_params(o.quadrature_rule, ke, ke, o.exterior, o.exterior, o.vacuum_permeability,
                                              o.vacuum_permittivity, omega, thr);
    host::MeshHandle mh(mesh);
    DeviceArray<std::complex<real_t>> Z((size_t)n * n), x(n);
    auto t0 = Clock::now();
    auto assemble = [&] {
        detail::check(sie_assemble_pec(mh.get(), &p, reinterpret_cast<double*>(Z.data()), n, nullptr));
    };
    assemble();
    const sie_plane_wave w = source.plane_wave();
    detail::check(sie_rhs_plane_waves(mh.get(), SIE_RHS_EFIE, o.quadrature_rule, &w, 1, (float)o.vacuum_permeability,
                                      (float)o.vacuum_permittivity, o.vacuum_permeability, o.vacuum_permittivity,
                                      reinterpret_cast<double*>(x.data()), n, nullptr));
    host::hip_ok(hipDeviceSynchronize(), "PEC assembly");
    mt.assembly_seconds = seconds_since(t0);
    if (o.capture_assembled_matrix) {
        result.assembled_matrix.resize((size_t)n * n);
        Z.download(result.assembled_matrix.data(), (size_t)n * n);
    }
    std::vector<std::complex<real_t>> b(n);
    x.download(b.data(), n);
    t0 = Clock::now();
    gpu::solve_dense_lu(n, Z.data(), x.data());
    mt.solve_seconds = seconds_since(t0);
    if (o.compute_residual) {
        t0 = Clock::now();
        assemble();
        residual(n, Z, x.data(), b, host::host_norm(b), mt.absolute_residual, mt.relative_residual);
        mt.residual_computed = true;
        mt.residual_seconds = seconds_since(t0);
    }
    result.exterior_current.resize(n);
    x.download(result.exterior_current.data(), n);
    for (const auto& v : result.exterior_current) {
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
            throw std::runtime_error("PEC solution contains a non-finite coefficient");
        mt.maximum_j_coefficient = std::max(mt.maximum_j_coefficient, std::hypot(v.real(), v.imag()));
    }
    return result;
}

DielectricBatchResult solve_dielectric_dense_batch(const mesh::MeshData& mesh, const DielectricOptions& o,
                                                   const std::vector<excitation::PreparedExcitation>& sources) {
    validate_common(mesh, o.frequency_hz, o.interior, o.exterior, o.vacuum_permittivity, o.vacuum_permeability,
                    o.quadrature_rule, o.singularity_threshold_lambda0);
    if (o.formulation != DielectricFormulation::T_PMCHWT)
        throw std::invalid_argument("only the T-PMCHWT dielectric formulation is provided by this library");
    if (sources.empty()) throw std::invalid_argument("at least one excitation is required");
    for (const auto& src : sources)
        src.validate_environment(o.frequency_hz, o.interior, o.exterior, o.vacuum_permittivity, o.vacuum_permeability);
    if (mesh.edges.size() > std::numeric_limits<std::size_t>::max() / 2) throw std::overflow_error("SIE matrix dimension overflows size_t");
    check_dimension(2 * mesh.edges.size());
    host::require_device();
    DielectricBatchResult batch;
    auto& mt = batch.metrics;
    mt.vertex_count = mesh.vertices.size();
    mt.face_count = mesh.faces.size();
    mt.edge_count = mesh.edges.size();
    mt.matrix_dimension = 2 * mesh.edges.size();
    mt.dense_matrix_bytes = mt.matrix_dimension * mt.matrix_dimension * sizeof(std::complex<real_t>);
    const int n = (int)mt.matrix_dimension;
    const int nrhs = (int)sources.size();
    const double omega = 2.0 * std::acos(-1.0) * o.frequency_hz;
    auto kof = [&](const IsotropicMedium& m) {
        return omega * std::sqrt(m.relative_permeability * o.vacuum_permeability * m.relative_permittivity *
                                 o.vacuum_permittivity);
    };
    const double thr = o.singularity_threshold_lambda0 * lambda0_of(omega, o.vacuum_permeability, o.vacuum_permittivity);
    const sie_calderon_params p = make_params(o.quadrature_rule, kof(o.interior), kof(o.exterior), o.interior,
                                              o.exterior, o.vacuum_permeability, o.vacuum_permittivity, omega, thr);
    auto t0 = Clock::now();
    host::MeshHandle mh(mesh);
    DeviceArray<std::complex<real_t>> Z((size_t)n * n), X((size_t)n * nrhs);
    mt.device_preparation_seconds = seconds_since(t0);
    t0 = Clock::now();
    auto assemble = [&] {
        detail::check(sie_assemble_pmchwt(mh.get(), &p, reinterpret_cast<double*>(Z.data()), n, nullptr));
    };
    assemble();
    std::vector<sie_plane_wave> waves;
    for (const auto& src : sources) waves.push_back(src.plane_wave());
    detail::check(sie_rhs_plane_waves(mh.get(), SIE_RHS_PMCHWT, o.quadrature_rule, waves.data(), nrhs,
                                      (float)o.vacuum_permeability, (float)o.vacuum_permittivity, o.vacuum_permeability,
                                      o.vacuum_permittivity, reinterpret_cast<double*>(X.data()), n, nullptr));
    host::hip_ok(hipDeviceSynchronize(), "PMCHWT assembly");
    mt.assembly_seconds = seconds_since(t0);
    std::vector<std::complex<real_t>> B((size_t)n * nrhs);
    X.download(B.data(), B.size());
    t0 = Clock::now();
    if (o.precision == SolvePrecision::fp64) {
        gpu::DenseLuSolver lu(n);
        lu.factorize(Z.data());
        lu.solve(gpu::Op::N, nrhs, Z.data(), X.data(), n);
    } else {
        gpu::MixedLuSolver mixed;
        mixed.factorize(Z.data(), n);
        DeviceArray<std::complex<real_t>> Bd((size_t)n * nrhs);
        Bd.upload(B.data(), B.size());
        const auto be = mixed.solve(gpu::Op::N, nrhs, Bd.data(), n, X.data(), n, o.refinement_tolerance,
                                    o.refinement_max_sweeps);
        for (double r : be)
            if (!(r <= o.refinement_tolerance))
                throw std::runtime_error("iterative refinement did not reach the requested backward error (" +
                                         std::to_string(r) + ")");
    }
    host::hip_ok(hipDeviceSynchronize(), "PMCHWT solve");
    mt.solve_seconds = seconds_since(t0);
    if (o.compute_residual) {
        t0 = Clock::now();
        assemble();
        mt.residual_computed = true;
    }
    const double eta0 = std::sqrt(o.vacuum_permeability / o.vacuum_permittivity);
    const std::size_t ne = mesh.edges.size();
    batch.cases.resize(nrhs);
    for (int c = 0; c < nrhs; ++c) {
        DielectricResult& r = batch.cases[c];
        r.metrics = mt;
        if (o.compute_residual) {
            const std::vector<std::complex<real_t>> b(B.begin() + (size_t)c * n, B.begin() + (size_t)(c + 1) * n);
            residual(n, Z, X.data() + (size_t)c * n, b, host::host_norm(b), r.metrics.absolute_residual,
                     r.metrics.relative_residual);
        }
        std::vector<std::complex<real_t>> cur(n);
        X.download(cur.data(), n, (size_t)c * n);
        for (std::size_t e = 0; e < ne; ++e) cur[ne + e] *= eta0;
        const auto mx = block_maxima(cur, ne, 2, "non-finite physical current coefficient at index ");
        r.metrics.maximum_j_coefficient = mx[0];
        r.metrics.maximum_m_coefficient = mx[1];
        r.physical_currents = std::move(cur);
    }
    if (o.compute_residual) mt.residual_seconds = seconds_since(t0);
    for (auto& r : batch.cases) r.metrics.residual_seconds = mt.residual_seconds;
    return batch;
}

DielectricResult solve_dielectric_dense(const mesh::MeshData& mesh, const DielectricOptions& o,
                                        const excitation::PreparedExcitation& source) {
    auto batch = solve_dielectric_dense_batch(mesh, o, {source});
    return std::move(batch.cases[0]);
}

}  // namespace solvers
}  // namespace sie_mom
