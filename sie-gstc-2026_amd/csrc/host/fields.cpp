// solvers::evaluate_surface_integral_fields over the C ABI: the reference's
// input checks (surface_current_fields.cu:46-90, 181-209) in the same order and
// wording, then sie_surface_fields on the device.
#include "sie_mom/solvers/surface_current_fields.hpp"

#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>

#include "host_common.hpp"

namespace sie_mom {
namespace solvers {

namespace {
bool finite(double v) { return std::isfinite(v); }
}  // namespace

std::vector<ElectromagneticFieldSample> evaluate_surface_integral_fields(
    const mesh::MeshData& mesh, const std::vector<std::complex<real_t>>& electric_currents,
    const std::vector<std::complex<real_t>>& magnetic_currents, const std::vector<vec3_t>& observation_points,
    real_t frequency_hz, const IsotropicMedium& medium, int quadrature_rule, real_t vacuum_permittivity,
    real_t vacuum_permeability) {
    if (!(frequency_hz > 0) || !finite(frequency_hz))
        throw std::invalid_argument("field-evaluation frequency must be positive and finite");
    if (!(medium.relative_permittivity > 0) || !(medium.relative_permeability > 0) ||
        !finite(medium.relative_permittivity) || !finite(medium.relative_permeability))
        throw std::invalid_argument("field-evaluation relative permittivity and permeability must be positive and finite");
    if (!(vacuum_permittivity > 0) || !(vacuum_permeability > 0) || !finite(vacuum_permittivity) ||
        !finite(vacuum_permeability))
        throw std::invalid_argument("field-evaluation vacuum constants must be positive and finite");
    if (quadrature_rule < 1 || quadrature_rule > 20)
        throw std::invalid_argument("Dunavant quadrature rule must be in [1, 20]");
    for (std::size_t i = 0; i < observation_points.size(); ++i) {
        const auto& p = observation_points[i];
        if (!finite(p.x) || !finite(p.y) || !finite(p.z))
            throw std::invalid_argument("non-finite observation point at index " + std::to_string(i));
    }
    if (observation_points.empty()) return {};
    if (mesh.vertices.empty() || mesh.faces.empty() || mesh.edges.empty())
        throw std::invalid_argument("field-evaluation mesh must contain vertices, faces, and edges");
    if (mesh.faces_edges.size() != mesh.faces.size())
        throw std::invalid_argument("field-evaluation face-to-edge mapping size does not match face count");
    if (mesh.faces.size() > static_cast<std::size_t>(std::numeric_limits<int>::max()))
        throw std::overflow_error("field-evaluation face count exceeds the kernel limit");
    if (electric_currents.size() != mesh.edges.size() || magnetic_currents.size() != mesh.edges.size())
        throw std::invalid_argument("field evaluation requires one electric and magnetic coefficient per mesh edge");
    host::require_device();
    auto check_finite = [](const std::vector<std::complex<real_t>>& v) {
        for (std::size_t i = 0; i < v.size(); ++i)
            if (!finite(v[i].real()) || !finite(v[i].imag()))
                throw std::invalid_argument("non-finite surface-current coefficient at index " + std::to_string(i));
    };
    check_finite(electric_currents);
    check_finite(magnetic_currents);

    const double pi = std::acos(-1.0);
    const double omega = 2.0 * pi * frequency_hz;
    const double eps = medium.relative_permittivity * vacuum_permittivity;
    const double mu = medium.relative_permeability * vacuum_permeability;

    host::MeshHandle handle(mesh);
    const std::size_t ne = mesh.edges.size(), np = observation_points.size();
    host::DeviceArray<std::complex<real_t>> J(ne), M(ne), E(3 * np), H(3 * np);
    host::DeviceArray<double> P(3 * np);
    J.upload(electric_currents.data(), ne);
    M.upload(magnetic_currents.data(), ne);
    std::vector<double> pts(3 * np);
    for (std::size_t i = 0; i < np; ++i) {
        pts[3 * i] = observation_points[i].x;
        pts[3 * i + 1] = observation_points[i].y;
        pts[3 * i + 2] = observation_points[i].z;
    }
    P.upload(pts.data(), pts.size());
    detail::check(sie_surface_fields(handle.get(), quadrature_rule, omega, eps, mu,
                                     reinterpret_cast<const double*>(J.data()), reinterpret_cast<const double*>(M.data()),
                                     P.data(), (int)np, reinterpret_cast<double*>(E.data()),
                                     reinterpret_cast<double*>(H.data()), nullptr));
    std::vector<std::complex<real_t>> he(3 * np), hh(3 * np);
    E.download(he.data(), he.size());
    H.download(hh.data(), hh.size());
    std::vector<ElectromagneticFieldSample> out(np);
    for (std::size_t i = 0; i < np; ++i)
        for (int c = 0; c < 3; ++c) {
            out[i].electric[c] = he[3 * i + c];
            out[i].magnetic[c] = hh[3 * i + c];
        }
    return out;
}

}  // namespace solvers
}  // namespace sie_mom
