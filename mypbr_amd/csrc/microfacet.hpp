// microfacet.hpp - GGX, Fresnel and Kulla-Conty building blocks, host and
// device. Operand order follows the reference's microfacet.h so that the
// host-built albedo table is bit-identical to the reference's and device
// evaluations agree to rounding.
#pragma once

#include "geom.hpp"

namespace mpt {

constexpr int kGgxTableSize = 32;

MPT_HD double roughness_to_alpha(double r) {  // microfacet.h:17-19
    double a = r * r;
    return a > 1e-3 ? a : 1e-3;
}

MPT_HD double pow5(double x) { return std::pow(x, 5); }

MPT_HD double fresnel_schlick(double cos_theta, double f0) {  // microfacet.h:22-24
    return f0 + (1 - f0) * pow5(1 - clampd(cos_theta, 0.0, 1.0));
}
MPT_HD v3 fresnel_schlick3(double cos_theta, const v3& f0) {
    return v3{fresnel_schlick(cos_theta, f0.x), fresnel_schlick(cos_theta, f0.y), fresnel_schlick(cos_theta, f0.z)};
}

MPT_HD double fresnel_dielectric(double cos_i, double eta) {  // microfacet.h:34-48
    cos_i = clampd(cos_i, -1.0, 1.0);
    if (cos_i < 0) {
        eta = 1 / eta;
        cos_i = -cos_i;
    }
    double sin2_i = 1 - cos_i * cos_i;
    double sin2_t = sin2_i / (eta * eta);
    if (sin2_t >= 1) return 1;
    double cos_t = std::sqrt(1 - sin2_t);
    double r_parl = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    double r_perp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return (r_parl * r_parl + r_perp * r_perp) / 2;
}

// Isotropic GGX with height-correlated Smith masking (microfacet.h:145-212).
MPT_HD double ggx_D(double alpha, const v3& h) {
    if (h.z <= 0) return 0;
    double a2 = alpha * alpha;
    double d = h.z * h.z * (a2 - 1) + 1;
    return a2 / (kPi * d * d);
}
MPT_HD double ggx_lambda(double alpha, const v3& w) {
    double cos2 = w.z * w.z;
    if (cos2 <= 0) return kInf;
    double tan2 = std::fmax(0.0, 1 - cos2) / cos2;
    return (-1 + std::sqrt(1 + alpha * alpha * tan2)) / 2;
}
MPT_HD double ggx_G1(double alpha, const v3& w) { return 1 / (1 + ggx_lambda(alpha, w)); }
MPT_HD double ggx_G(double alpha, const v3& wo, const v3& wi) {
    return 1 / (1 + ggx_lambda(alpha, wo) + ggx_lambda(alpha, wi));
}
MPT_HD double ggx_visible_pdf(double alpha, const v3& wo, const v3& h) {
    if (wo.z <= 0) return 0;
    return ggx_G1(alpha, wo) / wo.z * ggx_D(alpha, h) * std::fabs(dot(wo, h));
}
// Heitz 2018 visible-normal sampling driven by u1, u2 in [0, 1).
MPT_HD v3 ggx_sample_visible(double alpha, const v3& wo, double u1, double u2) {
    v3 vh = unit_vector(v3{alpha * wo.x, alpha * wo.y, wo.z});
    double lensq = vh.x * vh.x + vh.y * vh.y;
    v3 t1 = lensq > 0 ? v3{-vh.y, vh.x, 0} / std::sqrt(lensq) : v3{1, 0, 0};
    v3 t2 = cross(vh, t1);
    double r = std::sqrt(u1);
    double sp, cp;
    sin_cos_2pi(u2, sp, cp);  // phi = 2 pi u2
    double p1 = r * cp;
    double p2 = r * sp;
    double s = 0.5 * (1 + vh.z);
    p2 = (1 - s) * std::sqrt(std::fmax(0.0, 1 - p1 * p1)) + s * p2;
    v3 nh = p1 * t1 + p2 * t2 + std::sqrt(std::fmax(0.0, 1 - p1 * p1 - p2 * p2)) * vh;
    return unit_vector(v3{alpha * nh.x, alpha * nh.y, std::fmax(1e-6, nh.z)});
}

// Bilinear E(mu, alpha) and linear E_avg(alpha) from the 32x32 table
// (microfacet.h:276-293). e is [roughness row][mu column].
MPT_HD double ggx_E(const double* e, double cos_theta, double alpha) {
    const int n = kGgxTableSize;
    double x = clampd(cos_theta, 0.0, 1.0) * (n - 1);
    double y = clampd(std::sqrt(alpha), 0.0, 1.0) * (n - 1);
    int i = int(x) < n - 2 ? int(x) : n - 2;
    int j = int(y) < n - 2 ? int(y) : n - 2;
    double tx = x - i, ty = y - j;
    double lo = (1 - tx) * e[j * n + i] + tx * e[j * n + i + 1];
    double hi = (1 - tx) * e[(j + 1) * n + i] + tx * e[(j + 1) * n + i + 1];
    return (1 - ty) * lo + ty * hi;
}
MPT_HD double ggx_E_avg(const double* e_avg, double alpha) {
    const int n = kGgxTableSize;
    double y = clampd(std::sqrt(alpha), 0.0, 1.0) * (n - 1);
    int j = int(y) < n - 2 ? int(y) : n - 2;
    double ty = y - j;
    return (1 - ty) * e_avg[j] + ty * e_avg[j + 1];
}
// f_ms cos_i of the compensation lobe (microfacet.h:341-349).
MPT_HD double ggx_ms_eval(const double* e, const double* e_avg, double cos_o, double cos_i, double alpha) {
    if (cos_o <= 0 || cos_i <= 0) return 0;
    double missing_avg = 1 - ggx_E_avg(e_avg, alpha);
    if (missing_avg < 1e-6) return 0;
    return (1 - ggx_E(e, cos_o, alpha)) * (1 - ggx_E(e, cos_i, alpha)) / (kPi * missing_avg) * cos_i;
}
MPT_HD double ggx_ms_albedo(const double* e, const double* e_avg, double cos_o, double alpha) {
    if (cos_o <= 0 || 1 - ggx_E_avg(e_avg, alpha) < 1e-6) return 0;
    return 1 - ggx_E(e, cos_o, alpha);
}
// F_ms = F_avg^2 E_avg / (1 - F_avg (1 - E_avg)) (microfacet.h:363-369).
MPT_HD v3 ggx_ms_fresnel_avg(const double* e_avg, const v3& f_avg, double alpha) {
    double ea = ggx_E_avg(e_avg, alpha);
    return v3{f_avg.x * f_avg.x * ea / (1 - f_avg.x * (1 - ea)), f_avg.y * f_avg.y * ea / (1 - f_avg.y * (1 - ea)),
              f_avg.z * f_avg.z * ea / (1 - f_avg.z * (1 - ea))};
}
MPT_HD v3 ggx_ms_fresnel(const double* e_avg, const v3& f0, double alpha) {
    return ggx_ms_fresnel_avg(e_avg, f0 + (v3{1, 1, 1} - f0) / 21, alpha);
}

}  // namespace mpt
