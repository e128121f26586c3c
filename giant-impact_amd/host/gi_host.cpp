// gi_host.cpp -- units, kernel constants, parameter files and seeded
// synthetic bodies for the Giant-Impact step loop (host side, C ABI).
#include "../../include/giant_impact_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "gi_host_internal.h"

using namespace gi_host;


extern "C" {

void gi_generator_setup_defaults(gi_generator_setup* s, int n) {
    std::memset(s, 0, sizeof *s);
    // SetupTemplates/setupGeneratingTargetImpactor:3-65
    s->initial_spin1[0] = 0.0f;
    s->initial_spin1[1] = 1.0f;
    s->initial_spin1[2] = 0.0f;
    s->initial_spin1[3] = 0.43f;
    s->initial_spin2[0] = 0.0f;
    s->initial_spin2[1] = 1.0f;
    s->initial_spin2[2] = 0.0f;
    s->initial_spin2[3] = 0.43f;
    s->fraction_earth_mass_body1 = 0.55;
    s->fraction_earth_mass_body2 = 0.55;
    s->fraction_fe_body1 = 0.26;
    s->fraction_si_body1 = 0.74;
    s->fraction_fe_body2 = 0.26;
    s->fraction_si_body2 = 0.74;
    s->damp_rate_body1 = 8.0f;
    s->damp_rate_body2 = 8.0f;
    s->total_number_of_elements = n;
    s->damp_time = 20.0;
    s->damp_rest_time = 20.0;
    s->spin_rest_time = 30.0;
    s->dt = 0.002;
    s->density_fe = 7.850e12;
    s->density_si = 3.320e12;
    s->k_fe = 1.5e14;
    s->k_si = 1.5e14;
    s->kr_fe = 0.01;
    s->kr_si = 0.01;
    s->sd_fe = 0.02;
    s->sd_si = 0.002;
    s->draw_rate = 50;
}

int gi_generator_setup_read(const char* path, gi_generator_setup* s) {
    if (!path || !s) return GI_E_INVALID_ARG;
    PositionalReader r(path);
    if (!r.ok()) return GI_E_IO;
    bool ok = true;
    // Order of generateTargetImpactor.cu:177-248.
    for (int k = 0; k < 4; ++k) ok = ok && r.next(s->initial_spin1[k]);
    for (int k = 0; k < 4; ++k) ok = ok && r.next(s->initial_spin2[k]);
    ok = ok && r.next(s->fraction_earth_mass_body1) && r.next(s->fraction_earth_mass_body2);
    ok = ok && r.next(s->fraction_fe_body1) && r.next(s->fraction_si_body1);
    ok = ok && r.next(s->fraction_fe_body2) && r.next(s->fraction_si_body2);
    ok = ok && r.next(s->damp_rate_body1) && r.next(s->damp_rate_body2);
    ok = ok && r.next(s->total_number_of_elements);
    ok = ok && r.next(s->damp_time) && r.next(s->damp_rest_time) && r.next(s->spin_rest_time);
    ok = ok && r.next(s->dt);
    ok = ok && r.next(s->density_fe) && r.next(s->density_si);
    ok = ok && r.next(s->k_fe) && r.next(s->k_si);
    ok = ok && r.next(s->kr_fe) && r.next(s->kr_si);
    ok = ok && r.next(s->sd_fe) && r.next(s->sd_si);
    ok = ok && r.next(s->draw_rate);
    return ok ? 0 : GI_E_PARSE;
}

int gi_units_compute(const gi_generator_setup* s, gi_units* u) {
    if (!s || !u) return GI_E_INVALID_ARG;
    std::memset(u, 0, sizeof *u);
    const double m1 = kMassOfEarth * s->fraction_earth_mass_body1;
    const double m2 = kMassOfEarth * s->fraction_earth_mass_body2;
    if (s->fraction_fe_body1 + s->fraction_si_body1 != 1.0) return GI_E_INVALID_ARG;
    if (s->fraction_fe_body2 + s->fraction_si_body2 != 1.0) return GI_E_INVALID_ARG;
    const double fe1 = s->fraction_fe_body1 * m1, si1 = s->fraction_si_body1 * m1;
    const double fe2 = s->fraction_fe_body2 * m2, si2 = s->fraction_si_body2 * m2;
    const double fe = fe1 + fe2, si = si1 + si2;
    const int n = s->total_number_of_elements;
    const double rho_ratio = s->density_si / s->density_fe;
    // Element counts from equal element volumes (generateTargetImpactor.cu:319-331).
    u->nfe = fe != 0.0 ? (int)((double)n * rho_ratio / (si / fe + rho_ratio)) : 0;
    u->nsi = n - u->nfe;
    u->nfe1 = fe != 0.0 ? (int)(u->nfe * fe1 / fe) : 0;
    u->nfe2 = u->nfe - u->nfe1;
    u->nsi1 = si != 0.0 ? (int)(u->nsi * si1 / si) : 0;
    u->nsi2 = u->nsi - u->nsi1;
    const double mfe = u->nfe != 0 ? fe / u->nfe : 0.0;
    const double msi = u->nsi != 0 ? si / u->nsi : 0.0;
    // Length unit = element diameter, mass unit = silicate element, time unit
    // makes G = 1 (:340-358).
    if (u->nsi != 0) {
        u->unit_length = std::pow((6.0 * msi) / (kPi * s->density_si), (1.0 / 3.0));
        u->unit_mass = msi;
        u->unit_time = std::sqrt((6.0 * msi * (double)u->nsi) / (kUniversalGravity * kPi * s->density_si * si));
        u->mass_si = 1.0;
        u->mass_fe = s->density_fe / s->density_si;
    } else if (u->nfe != 0) {
        u->unit_length = std::pow((6.0 * mfe) / (kPi * s->density_fe), (1.0 / 3.0));
        u->unit_mass = mfe;
        u->unit_time = std::sqrt((6.0 * mfe * (double)u->nfe) / (kUniversalGravity * kPi * s->density_fe * fe));
        u->mass_fe = 1.0;
        u->mass_si = -1.0;
    } else {
        return GI_E_INVALID_ARG;
    }
    u->diameter = 1.0;
    u->gravity = 1.0;
    u->mass_of_body1 = m1 / u->unit_mass;
    u->mass_of_body2 = m2 / u->unit_mass;
    u->spin1_w = (float)(s->initial_spin1[3] * (u->unit_time / 3600.0));
    u->spin2_w = (float)(s->initial_spin2[3] * (u->unit_time / 3600.0));
    u->damp_time = s->damp_time * (3600.0 / u->unit_time);
    u->damp_rest_time = s->damp_rest_time * (3600.0 / u->unit_time);
    u->spin_rest_time = s->spin_rest_time * (3600.0 / u->unit_time);
    const double kscale = u->unit_time * u->unit_time * u->unit_length / u->unit_mass;
    u->k_fe = s->k_fe * kscale;
    u->k_si = s->k_si * kscale;
    u->kr_fe = s->kr_fe;
    u->kr_si = s->kr_si;
    u->sd_fe = s->sd_fe;
    u->sd_si = s->sd_si;
    return 0;
}

// The pair-law constants common to both kernel families.
static void fill_law(double gravity, double mass_fe, double mass_si, double diameter, double kfe,
                     double ksi, double krfe, double krsi, double sdfe, double sdsi, float* out13) {
    out13[0] = (float)(gravity * mass_fe * mass_fe);  // GMassFeFe
    out13[1] = (float)(gravity * mass_fe * mass_si);  // GMassFeSi
    out13[2] = (float)(2.0 * kfe);                    // KFeFe
    out13[3] = (float)(2.0 * ksi);                    // KSiSi
    out13[4] = (float)(kfe + ksi);                    // KFeSi
    out13[5] = (float)(2.0 * kfe * krfe);             // KRFeFe
    out13[6] = (float)(2.0 * ksi * krsi);             // KRSiSi
    out13[7] = (float)(kfe * krfe + ksi * krsi);      // KRFeSi
    out13[8] = (float)(sdfe >= sdsi ? kfe + ksi * krsi : kfe * krfe + ksi);  // KRMix
    out13[9] = (float)(diameter - diameter * sdfe);   // ShellBreakFe
    out13[10] = (float)(diameter - diameter * sdsi);  // ShellBreakSi
    // ShellBreakFeSi1 is the outer (thinner) shell, FeSi2 the inner one.
    out13[11] = (float)(diameter - diameter * std::min(sdfe, sdsi));
    out13[12] = (float)(diameter - diameter * std::max(sdfe, sdsi));
}

void gi_separate_constants(const gi_units* u, double dt, gi_force_separate_constants* fc,
                           gi_move_separate_constants* mc) {
    if (fc) {
        fill_law(u->gravity, u->mass_fe, u->mass_si, u->diameter, u->k_fe, u->k_si, u->kr_fe,
                 u->kr_si, u->sd_fe, u->sd_si, &fc->GMassFeFe);
        fc->boarder1 = u->nfe1;
        fc->boarder2 = u->nfe1 + u->nsi1;
        fc->boarder3 = u->nfe1 + u->nsi1 + u->nfe2;
    }
    if (mc) {
        mc->Dt = (float)dt;
        mc->DtOverMassFe = (float)(dt / u->mass_fe);
        mc->DtOverMassSi = (float)(dt / u->mass_si);
        mc->boarder1 = u->nfe1;
        mc->boarder2 = u->nsi1 + u->nfe1;
        mc->boarder3 = u->nfe1 + u->nsi1 + u->nfe2;
    }
}

void gi_impact_constants(const gi_units* u, double dt, gi_force_collision_constants* fc,
                         gi_move_collision_constants* mc) {
    // The impact reads MassFe .. SDSi back from a "%f" text file, so every one
    // of them carries 6 decimals; Gravity is the literal 1 (impactAddTime.cu:131-132).
    const double mfe = round_trip_f(u->mass_fe), msi = round_trip_f(u->mass_si);
    const double dia = round_trip_f(u->diameter);
    const double kfe = round_trip_f(u->k_fe), ksi = round_trip_f(u->k_si);
    const double krfe = round_trip_f(u->kr_fe), krsi = round_trip_f(u->kr_si);
    const double sdfe = round_trip_f(u->sd_fe), sdsi = round_trip_f(u->sd_si);
    const int nfe = u->nfe1 + u->nfe2;
    if (fc) {
        fill_law(1.0, mfe, 1.0, dia, kfe, ksi, krfe, krsi, sdfe, sdsi, &fc->GMassFeFe);
        fc->NFe = nfe;
    }
    if (mc) {
        mc->Dt = (float)dt;
        mc->DtOverMassSi = (float)(dt / msi);
        mc->DtOverMassFe = (float)(dt / mfe);
        mc->NFe = nfe;
    }
}

int gi_parameters_write(const char* path, const gi_units* u, int total_number_of_elements) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return GI_E_IO;
    // Field order and format of generateTargetImpactorBasicFunctions.h:693-709.
    std::fprintf(f, "\n UnitTime = %f", u->unit_time);
    std::fprintf(f, "\n UnitLength = %f", u->unit_length);
    std::fprintf(f, "\n UnitMass = %f", u->unit_mass);
    std::fprintf(f, "\n MassFe = %f", u->mass_fe);
    std::fprintf(f, "\n MassSi = %f", u->mass_si);
    std::fprintf(f, "\n Diameter = %f", u->diameter);
    std::fprintf(f, "\n KFe = %f", u->k_fe);
    std::fprintf(f, "\n KSi = %f", u->k_si);
    std::fprintf(f, "\n KRFe = %f", u->kr_fe);
    std::fprintf(f, "\n KRSi = %f", u->kr_si);
    std::fprintf(f, "\n SDFe = %f", u->sd_fe);
    std::fprintf(f, "\n SDSi = %f", u->sd_si);
    std::fprintf(f, "\n TotalNumberOfElements = %d", total_number_of_elements);
    std::fprintf(f, "\n NFe1 = %d", u->nfe1);
    std::fprintf(f, "\n NSi1 = %d", u->nsi1);
    std::fprintf(f, "\n NFe2 = %d", u->nfe2);
    std::fprintf(f, "\n NSi2 = %d", u->nsi2);
    return std::fclose(f) == 0 ? 0 : GI_E_IO;
}

int gi_parameters_read(const char* path, gi_units* u, int* total_number_of_elements) {
    if (!path || !u || !total_number_of_elements) return GI_E_INVALID_ARG;
    PositionalReader r(path);
    if (!r.ok()) return GI_E_IO;
    std::memset(u, 0, sizeof *u);
    bool ok = r.next(u->unit_time) && r.next(u->unit_length) && r.next(u->unit_mass) &&
              r.next(u->mass_fe) && r.next(u->mass_si) && r.next(u->diameter) && r.next(u->k_fe) &&
              r.next(u->k_si) && r.next(u->kr_fe) && r.next(u->kr_si) && r.next(u->sd_fe) &&
              r.next(u->sd_si) && r.next(*total_number_of_elements) && r.next(u->nfe1) &&
              r.next(u->nsi1) && r.next(u->nfe2) && r.next(u->nsi2);
    if (!ok) return GI_E_PARSE;
    u->nfe = u->nfe1 + u->nfe2;  // impactBasicFunctions.h:105-106
    u->nsi = u->nsi1 + u->nsi2;
    u->gravity = 1.0;
    return 0;
}

int gi_make_impact(const gi_impact_config* cfg, gi_float4* pos, gi_float4* vel,
                   gi_force_collision_constants* fc, gi_move_collision_constants* mc,
                   gi_impact_info* info) {
    if (!cfg || !pos || !vel || cfg->n_target <= 0 || cfg->n_impactor < 0 || cfg->fe_fraction < 0.0 ||
        cfg->fe_fraction > 1.0)
        return GI_E_INVALID_ARG;
    const int nt = cfg->n_target, ni = cfg->n_impactor, n = nt + ni;
    gi_generator_setup setup;
    gi_generator_setup_defaults(&setup, n);
    gi_units u;
    int rc = gi_units_compute(&setup, &u);
    if (rc) return rc;
    gi_force_collision_constants f;
    gi_move_collision_constants m;
    gi_impact_constants(&u, setup.dt, &f, &m);
    const double mfe = f.GMassFeSi, msi = 1.0;

    SplitMix64 rng{cfg->seed};
    Body bt = make_body(nt, cfg->fe_fraction, cfg->jitter, rng);
    Body bi = ni > 0 ? make_body(ni, cfg->fe_fraction, cfg->jitter, rng) : Body();

    auto body_mass = [&](const Body& b, int nb) { return b.nfe * mfe + (nb - b.nfe) * msi; };
    const double Mt = body_mass(bt, nt), Mi = ni > 0 ? body_mass(bi, ni) : 0.0;
    // Volume-equivalent radius of nb unit cells.
    const double Rt = std::cbrt(3.0 * nt / (4.0 * kPi));
    const double Ri = ni > 0 ? std::cbrt(3.0 * ni / (4.0 * kPi)) : 0.0;
    const double gap = (cfg->gap_in_target_radii > 0.0 ? cfg->gap_in_target_radii : 1.0) * Rt;
    const double vesc = std::sqrt(2.0 * (Mt + Mi) / (Rt + Ri));
    const double vrel = cfg->speed_vesc * vesc;
    const double theta = cfg->angle_deg * kPi / 180.0;
    const double b = (Rt + Ri) * std::sin(theta);
    const double d = Rt + Ri + gap;
    const double relx = -std::sqrt(std::max(d * d - b * b, 0.0)), rely = b;
    const double M = Mt + Mi;
    // Body centres and velocities in the centre-of-mass frame.
    const double ct[3] = {-Mi / M * relx, -Mi / M * rely, 0.0};
    const double ci[3] = {Mt / M * relx, Mt / M * rely, 0.0};
    const double vt[3] = {-Mi / M * vrel, 0.0, 0.0};
    const double vi[3] = {Mt / M * vrel, 0.0, 0.0};
    const double wt = cfg->target_spin * std::sqrt(Mt / (Rt * Rt * Rt));
    const double wi = ni > 0 ? cfg->impactor_spin * std::sqrt(Mi / (Ri * Ri * Ri)) : 0.0;

    // Rigid spin about +z through each body's own centre of mass:
    // v += omega * z_hat x r (the velocity spinBody adds,
    // generateTargetImpactorBasicFunctions.h:442-468).
    auto body_com = [&](const Body& bb, int nb) {
        double c[3] = {0, 0, 0}, mt = 0;
        for (int k = 0; k < nb; ++k) {
            double mk = k < bb.nfe ? mfe : msi;
            c[0] += mk * bb.x[k];
            c[1] += mk * bb.y[k];
            c[2] += mk * bb.z[k];
            mt += mk;
        }
        return std::vector<double>{c[0] / mt, c[1] / mt, c[2] / mt};
    };
    std::vector<double> comt = body_com(bt, nt);
    std::vector<double> comi = ni > 0 ? body_com(bi, ni) : std::vector<double>{0, 0, 0};

    // Layout [Fe1, Fe2, Si1, Si2] (impactInitialization.cu:329-371).
    int k = 0;
    auto emit = [&](const Body& bb, int lo, int hi, const double* c, const double* v, const std::vector<double>& com,
                    double w, float tag, float mass) {
        for (int q = lo; q < hi; ++q, ++k) {
            const double rx = bb.x[q] - com[0], ry = bb.y[q] - com[1];
            pos[k].x = (float)(c[0] + rx);
            pos[k].y = (float)(c[1] + ry);
            pos[k].z = (float)(c[2] + bb.z[q] - com[2]);
            pos[k].w = tag;
            vel[k].x = (float)(v[0] - w * ry);
            vel[k].y = (float)(v[1] + w * rx);
            vel[k].z = (float)v[2];
            vel[k].w = mass;
        }
    };
    emit(bt, 0, bt.nfe, ct, vt, comt, wt, 0.0f, (float)mfe);
    if (ni > 0) emit(bi, 0, bi.nfe, ci, vi, comi, wi, 2.0f, (float)mfe);
    emit(bt, bt.nfe, nt, ct, vt, comt, wt, 1.0f, (float)msi);
    if (ni > 0) emit(bi, bi.nfe, ni, ci, vi, comi, wi, 3.0f, (float)msi);

    f.NFe = bt.nfe + (ni > 0 ? bi.nfe : 0);
    m.NFe = f.NFe;
    if (fc) *fc = f;
    if (mc) *mc = m;
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->n = n;
        info->nfe = f.NFe;
        info->nfe1 = bt.nfe;
        info->nfe2 = ni > 0 ? bi.nfe : 0;
        info->nsi1 = nt - bt.nfe;
        info->nsi2 = ni > 0 ? ni - bi.nfe : 0;
        info->mass_fe = mfe;
        info->mass_si = msi;
        info->radius_target = Rt;
        info->radius_impactor = Ri;
        info->mass_target = Mt;
        info->mass_impactor = Mi;
        info->v_escape = vesc;
        info->v_relative = vrel;
        info->impact_parameter = b;
        info->target_omega = wt;
        info->impactor_omega = wi;
        double mt = 0;
        for (int q = 0; q < n; ++q) {
            double mq = q < f.NFe ? mfe : msi;
            info->com[0] += mq * pos[q].x;
            info->com[1] += mq * pos[q].y;
            info->com[2] += mq * pos[q].z;
            info->vel[0] += mq * vel[q].x;
            info->vel[1] += mq * vel[q].y;
            info->vel[2] += mq * vel[q].z;
            mt += mq;
        }
        for (int a = 0; a < 3; ++a) {
            info->com[a] /= mt;
            info->vel[a] /= mt;
        }
    }
    return 0;
}

}  // extern "C"
