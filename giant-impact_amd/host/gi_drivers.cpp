// gi_drivers.cpp -- the reference's three compute programs without their
// display: body generation (generateTargetImpactor.exe), impact set-up
// (impactInitialization.exe) and the impact run (impactAddTime.exe), on the
// device-resident steppers of this library and the reference's file formats.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <limits>
#include <vector>

#include "../../include/giant_impact_host.h"
#include "gi_host_internal.h"

using namespace gi_host;

namespace {

struct DeviceState {
    void *pos = nullptr, *vel = nullptr, *force = nullptr;
    int alloc(int n) {
        const size_t b = (size_t)(n > 0 ? n : 1) * 16;
        int rc;
        if ((rc = gi_malloc(&pos, b)) || (rc = gi_malloc(&vel, b)) || (rc = gi_malloc(&force, b))) return rc;
        return 0;
    }
    ~DeviceState() {
        if (pos) gi_free(pos);
        if (vel) gi_free(vel);
        if (force) gi_free(force);
    }
};

std::string join(const char *dir, const char *name) {
    std::string p(dir ? dir : ".");
    if (!p.empty() && p.back() != '/') p += '/';
    return p + name;
}

// Mass-weighted mean of v over a body (iron [a, b), silicate [b, c)) divided
// by the body's nominal mass, as getCenterOfMass / getAverageLinearVelocity
// do (generateTargetImpactorBasicFunctions.h:29-80,119-171).
void body_mean(const gi_float4 *v, int a, int b, int c, double mass_fe, double mass_si, double total,
               double out[3]) {
    out[0] = out[1] = out[2] = 0.0;
    if (total < 0.00000000001) return;
    for (int i = a; i < b; ++i) {
        out[0] += v[i].x * mass_fe;
        out[1] += v[i].y * mass_fe;
        out[2] += v[i].z * mass_fe;
    }
    for (int i = b; i < c; ++i) {
        out[0] += v[i].x * mass_si;
        out[1] += v[i].y * mass_si;
        out[2] += v[i].z * mass_si;
    }
    out[0] /= total;
    out[1] /= total;
    out[2] /= total;
}

int write_pos_vel(FILE *f, const gi_float4 *pos, const gi_float4 *vel, int n) {
    if (std::fwrite(pos, sizeof(gi_float4), n, f) != (size_t)n) return GI_E_IO;
    if (std::fwrite(vel, sizeof(gi_float4), n, f) != (size_t)n) return GI_E_IO;
    return 0;
}

int write_centering(const std::string &path, const float c[6]) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return GI_E_IO;
    // Field names and format of impactInitialization.cu:523-531.
    static const char *names[6] = {"InitialCenterOfMassX",   "InitialCenterOfMassY",   "InitialCenterOfMassZ",
                                   "InitialLinearVelocityX", "InitialLinearVelocityY", "InitialLinearVelocityZ"};
    for (int k = 0; k < 6; ++k) std::fprintf(f, "\n %s = %f\n", names[k], c[k]);
    return std::fclose(f) == 0 ? 0 : GI_E_IO;
}

// initialStatsFile: the text, order and unit conversions of
// recordInitialImpactStat (impactInitialization.cu:477-513).  mass, com, vel
// and L are the whole system's, in simulation units.
int write_initial_stats(const std::string &path, const gi_units &u, double mass, const double com[3],
                        const double vel[3], const double L[3]) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return GI_E_IO;
    const double mass_c = u.unit_mass, length_c = u.unit_length, velocity_c = u.unit_length / u.unit_time;
    const double momentum_c = u.unit_mass * u.unit_length * u.unit_length / u.unit_time;
    std::fprintf(f, "\n\n\n*************************************************************************\n\n");
    std::fprintf(f, "\nThe following are the initial stats of the impact\n");
    std::fprintf(f, "\nDistance is measured in Kilometers");
    std::fprintf(f, "\nMass is measured in Kilograms");
    std::fprintf(f, "\nTime is measured in seconds");
    std::fprintf(f, "\nVelocity is measured in Kilometers/second");
    std::fprintf(f, "\nAngular momentun is measured in Kilograms*Kilometers*Kilometers/seconds\n");
    std::fprintf(f, "\nCenter of mass of the entire system \t\t        = (%f, %f, %f)\n", com[0] * length_c,
                 com[1] * length_c, com[2] * length_c);
    std::fprintf(f, "\nTotal mass of the entire system \t= %e\n", mass * mass_c);
    double x = vel[0] * velocity_c, y = vel[1] * velocity_c, z = vel[2] * velocity_c;
    std::fprintf(f, "\nLinear velocity of the entire system system \t\t= (%f, %f, %f)", x, y, z);
    std::fprintf(f, "\nMagnitude of the linear velocity of the entire system \t= %f\n", std::sqrt(x * x + y * y + z * z));
    x = L[0] * momentum_c;
    y = L[1] * momentum_c;
    z = L[2] * momentum_c;
    std::fprintf(f, "\nAngular momentum of the entire system system \t\t= (%e, %e, %e)", x, y, z);
    std::fprintf(f, "\nMagnitude of the angular momentum of the entire system \t= %e\n", std::sqrt(x * x + y * y + z * z));
    std::fprintf(f, "\n\n*************************************************************************\n\n\n");
    return std::fclose(f) == 0 ? 0 : GI_E_IO;
}

// getAngularMomentum of a body (iron [a, b), silicate [b, c)): about its
// centre of mass and relative to its mean velocity, both divided by the
// nominal body mass (generateTargetImpactorBasicFunctions.h:173-237).
void body_angmom(const gi_float4 *pos, const gi_float4 *vel, int a, int b, int c, double mass_fe, double mass_si,
                 double total, double out[3]) {
    double cm[3], um[3];
    body_mean(pos, a, b, c, mass_fe, mass_si, total, cm);
    body_mean(vel, a, b, c, mass_fe, mass_si, total, um);
    out[0] = out[1] = out[2] = 0.0;
    for (int i = a; i < c; ++i) {
        const double m = i < b ? mass_fe : mass_si;
        const double rx = pos[i].x - cm[0], ry = pos[i].y - cm[1], rz = pos[i].z - cm[2];
        const double vx = vel[i].x - um[0], vy = vel[i].y - um[1], vz = vel[i].z - um[2];
        out[0] += (ry * vz - rz * vy) * m;
        out[1] += -(rx * vz - rz * vx) * m;
        out[2] += (rx * vy - ry * vx) * m;
    }
}

// targetImpactorStats: the text, order and conversions of
// recordTargetImpactorStats (generateTargetImpactorBasicFunctions.h:590-686)
// for the final bodies (layout [Fe1, Si1, Fe2, Si2]).  The body radius is the
// largest distance of an element from the body's centre of mass (the
// reference's getRadiusOfBody starts that maximum from an uninitialised
// value, :87; here from 0).
int write_target_impactor_stats(const std::string &path, const gi_generator_setup &s, const gi_units &u,
                                const gi_float4 *pos, const gi_float4 *vel) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return GI_E_IO;
    const double length_c = u.unit_length, mass_c = u.unit_mass, velocity_c = u.unit_length / u.unit_time;
    const double momentum_c = (u.unit_mass * u.unit_length * u.unit_length) / (u.unit_time);
    std::fprintf(f, "The conversion parameters to take you to and from our units to the real world units follow\n");
    std::fprintf(f, "\n Simulation length unit is %f kilometers", u.unit_length);
    std::fprintf(f, "\n Simulation mass unit is %e kilograms", u.unit_mass);
    std::fprintf(f, "\n Simulation time unit is %f seconds", u.unit_time);
    std::fprintf(f, "\n\n The initail statistics for this run in Simulation units follow\n");
    std::fprintf(f, "\n Diameter of an element: \t\t\tDiameter = %f", u.diameter);
    std::fprintf(f, "\n Gravitational constant: \t\t\tGravity = %f", u.gravity);
    std::fprintf(f, "\n The mass of a silicate element: \tMassSi = %f", u.mass_si);
    std::fprintf(f, "\n The mass of an iron element: \t\tMassFe = %f", u.mass_fe);
    std::fprintf(f, "\n\n The push back strength of iron: \tKFe = %f", u.k_fe);
    std::fprintf(f, "\n The push back strength of silicate: \tKSi = %f\n", u.k_si);
    std::fprintf(f, "\n The mass of body one: \tMassOfBody1 = %f", u.mass_of_body1);
    std::fprintf(f, "\n The mass of body two: \tMassOfBody2 = %f\n", u.mass_of_body2);
    const float *sp[2] = {s.initial_spin1, s.initial_spin2};
    const float sw[2] = {u.spin1_w, u.spin2_w};
    static const char *which[2] = {"one", "two"};
    for (int b = 0; b < 2; ++b) {
        // the direction's norm from float products, as the float4 fields give it
        const float q = sp[b][0] * sp[b][0] + sp[b][1] * sp[b][1] + sp[b][2] * sp[b][2];
        const double mag = std::sqrt((double)q);
        std::fprintf(f, "\n The initial spin in revolutions per time unit of body %s: (%f, %f, %f, %f)%s", which[b],
                     sp[b][0] / mag, sp[b][1] / mag, sp[b][2] / mag, (double)sw[b], b == 1 ? "\n" : "");
    }
    const int nfe = u.nfe1 + u.nfe2, nsi = u.nsi1 + u.nsi2;
    std::fprintf(f, "\n Total number of elements: \t\t\t\t\t\tTotalNumberOfElements = %d", nfe + nsi);
    std::fprintf(f, "\n Total number of iron elements: \t\t\t\t\tNFe = %d", nfe);
    std::fprintf(f, "\n Total number of silicate elements: \t\t\t\tNSi = %d", nsi);
    std::fprintf(f, "\n Total number of iron elements in body1: \t\t\tNFe1 = %d", u.nfe1);
    std::fprintf(f, "\n Total number of silicate elements in body1: \t\tNSi1 = %d", u.nsi1);
    std::fprintf(f, "\n Total number of iron elements in body2 \t\t\tNFe2: = %d", u.nfe2);
    std::fprintf(f, "\n Total number of silicate elements in body2: \t\tNSi2 = %d\n", u.nsi2);
    std::fprintf(f, "\n Time step in Simulation units: \t\tDt = %f", s.dt);
    std::fprintf(f, "\n Time step in seconds:\t\t\t\t\tDt = %f", s.dt * u.unit_time);
    std::fprintf(f, "\n Damp time in Simulation units: \t\tDampTime = %f", u.damp_time);
    std::fprintf(f, "\n Damp rest time in Simulation units: \tDampRestTime = %f", u.damp_rest_time);
    std::fprintf(f, "\n Spin rest time in Simulation units: \tSpinRestTime = %f", u.spin_rest_time);
    std::fprintf(f, "\n\n\n*****************************************************************************************************");
    std::fprintf(f, "\nThe follow are the statistics of the final created bodies in real world units");
    std::fprintf(f, "\n*****************************************************************************************************");
    const int b1 = u.nfe1, b2 = u.nfe1 + u.nsi1, b3 = b2 + u.nfe2, n = b3 + u.nsi2;
    const int ranges[2][3] = {{0, b1, b2}, {b2, b3, n}};
    const double nominal[2] = {u.mass_of_body1, u.mass_of_body2};
    const int counts[2][2] = {{u.nfe1, u.nsi1}, {u.nfe2, u.nsi2}};
    for (int b = 0; b < 2; ++b) {
        const int a0 = ranges[b][0], a1 = ranges[b][1], a2 = ranges[b][2];
        double com[3], v[3], L[3];
        body_mean(pos, a0, a1, a2, u.mass_fe, u.mass_si, nominal[b], com);
        double radius = 0.0;
        for (int i = a0; i < a2; ++i) {
            const double rx = (double)pos[i].x - com[0], ry = (double)pos[i].y - com[1], rz = (double)pos[i].z - com[2];
            radius = std::max(radius, std::sqrt(rx * rx + ry * ry + rz * rz));
        }
        const double mass = u.mass_fe * (double)counts[b][0] + u.mass_si * (double)counts[b][1];
        std::fprintf(f, "\n\n ***** Stats for Body%d *****", b + 1);
        std::fprintf(f, "\n Mass =  %e Kilograms", mass * mass_c);
        std::fprintf(f, "\n Radius =  %f Kilometers", radius * length_c);
        std::fprintf(f, "\n The center of mass = (%f, %f, %f) Kilometers from (0, 0, 0)", com[0] * length_c,
                     com[1] * length_c, com[2] * length_c);
        body_mean(vel, a0, a1, a2, u.mass_fe, u.mass_si, nominal[b], v);
        std::fprintf(f, "\n The average linear velocity = (%f, %f, %f) Kilometers/second", v[0] * velocity_c,
                     v[1] * velocity_c, v[2] * velocity_c);
        std::fprintf(f, "\n The magitude of the avergae linear velocity = %f Kilometers/second",
                     std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * velocity_c);
        body_angmom(pos, vel, a0, a1, a2, u.mass_fe, u.mass_si, nominal[b], L);
        std::fprintf(f, "\n The angular momentum = (%e, %e, %e) Kilograms*kilometers*kilometers/second",
                     L[0] * momentum_c, L[1] * momentum_c, L[2] * momentum_c);
        std::fprintf(f, "\n The magitude of the angular momentum = %e Kilograms*kilometers*kilometers/second",
                     std::sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]) * momentum_c);
    }
    std::fprintf(f, "\n\n");
    return std::fclose(f) == 0 ? 0 : GI_E_IO;
}

// The reference's "angle off the ecliptic plane" of a vector (the ecliptic is
// the xz-plane; impactBasicFunctions.h:976-980).
double off_ecliptic_deg(double x, double y, double z) {
    const double size = std::sqrt(x * x + y * y + z * z) * std::sqrt(x * x + z * z);
    const double angle = std::acos((x * x + z * z) / size);
    return 90.0 - angle * 180.0 / kPi;
}

}  // namespace

extern "C" {

int gi_write_run_stats(const char *path, const gi_earth_moon *em, const gi_units *u, double time) {
    if (!path || !em || !u) return GI_E_INVALID_ARG;
    FILE *f = std::fopen(path, "wb");
    if (!f) return GI_E_IO;
    // Text, order and unit conversions of recordFinalImpactStat
    // (impactBasicFunctions.h:883-1072).  Scopes: 0 system, 1 Earth-Moon,
    // 2 Earth, 3 Moon.
    const double time_c = u->unit_time, length_c = u->unit_length, mass_c = u->unit_mass;
    const double velocity_c = u->unit_length / u->unit_time;
    const double momentum_c = u->unit_mass * u->unit_length * u->unit_length / u->unit_time;
    std::fprintf(f, "\n\n\n*************************************************************************\n\n");
    std::fprintf(f, "\nThe following are the final stats of the run when time = %f hours\n", time * time_c / 3600.0);
    std::fprintf(f, "\nDistance is measured in Kilometers");
    std::fprintf(f, "\nMass is measured in Kilograms");
    std::fprintf(f, "\nTime is measured in seconds");
    std::fprintf(f, "\nVelocity is measured in Kilometers/second");
    std::fprintf(f, "\nAngular momentun is measured in Kilograms*Kilometers*Kilometers/seconds\n");
    std::fprintf(f, "\nThe mass of Earth \t\t= %e", em->mass[2] * mass_c);
    std::fprintf(f, "\nThe mass of Moon \t\t= %e", em->mass[3] * mass_c);
    if (em->mass[3] != 0.0) std::fprintf(f, "\nThe mass ratio Earth/Moon \t= %f\n", em->mass[2] / em->mass[3]);
    std::fprintf(f, "\nMoon iron from body 1 \t\t= %d", em->moon_fe1);
    std::fprintf(f, "\nMoon silicate from body 1 \t= %d", em->moon_si1);
    std::fprintf(f, "\nMoon iron from body 2 \t\t= %d", em->moon_fe2);
    std::fprintf(f, "\nMoon silicate from body 2 \t= %d", em->moon_si2);
    if (em->moon_fe2 + em->moon_si2 == 0)
        std::fprintf(f, "\nThe Moon is only composed of elements from body 1\n");
    else if (em->moon_fe1 + em->moon_si1 == 0)
        std::fprintf(f, "\nThe Moon is only composed of elements from body 2\n");
    else
        std::fprintf(f, "\nMoon ratio body1/body2 \t\t= %f\n",
                     (float)(em->moon_fe1 + em->moon_si1) / (float)(em->moon_fe2 + em->moon_si2));
    std::fprintf(f, "\nEarth iron from body 1 \t\t= %d", em->earth_fe1);
    std::fprintf(f, "\nEarth silicate from body 1 \t= %d", em->earth_si1);
    std::fprintf(f, "\nEarth iron from body 2 \t\t= %d", em->earth_fe2);
    std::fprintf(f, "\nEarth silicate from body 2 \t= %d", em->earth_si2);
    if (em->earth_fe2 + em->earth_si2 == 0)
        std::fprintf(f, "\nThe Earth is only composed of elements from body 1\n");
    else if (em->earth_fe1 + em->earth_si1 == 0)
        std::fprintf(f, "\nThe Earth is only composed of elements from body 2\n");
    else
        std::fprintf(f, "\nEarth ratio body1/body2 \t\t= %f\n",
                     (float)(em->earth_fe1 + em->earth_si1) / (float)(em->earth_fe2 + em->earth_si2));
    static const char *lname[3] = {"the Earth Moon system \t\t", "the Earth \t\t\t\t", "the Moon \t\t\t\t"};
    static const char *mname[3] = {"the system \t", "the Earth \t\t", "the Moon \t\t"};
    static const char *aname[3] = {"the system's", "the Earth's", "the Moon's"};
    for (int k = 0; k < 3; ++k) {
        const double x = em->angmom[k + 1][0] * momentum_c, y = em->angmom[k + 1][1] * momentum_c,
                     z = em->angmom[k + 1][2] * momentum_c;
        std::fprintf(f, "\nAngular momentum of %s= (%e, %e, %e)", lname[k], x, y, z);
        std::fprintf(f, "\nMagnitude of the angular momentum of %s= %e", mname[k], std::sqrt(x * x + y * y + z * z));
        std::fprintf(f, "\nAngle off ecliptic plane of %s rotation \t= %f\n", aname[k], off_ecliptic_deg(x, y, z));
    }
    static const char *cname[3] = {"the Earth-Moon system \t\t", "the Earth system \t\t\t",
                                   "the Moon system \t\t\t"};
    for (int k = 0; k < 3; ++k)
        std::fprintf(f, "\nCenter of mass of %s= (%f, %f, %f)%s", cname[k], em->com[k + 1][0] * length_c,
                     em->com[k + 1][1] * length_c, em->com[k + 1][2] * length_c, k == 2 ? "\n" : "");
    for (int k = 0; k < 3; ++k)
        std::fprintf(f, "\nLinear Velocity of %s= (%f, %f, %f)%s", cname[k], em->vel[k + 1][0] * velocity_c,
                     em->vel[k + 1][1] * velocity_c, em->vel[k + 1][2] * velocity_c, k == 2 ? "\n" : "");
    std::fprintf(f, "\n*****Stats of the entire system to check the numerical scheme's validity*****\n");
    std::fprintf(f, "\nCenter of mass of the entire system \t\t        = (%f, %f, %f)\n", em->com[0][0] * length_c,
                 em->com[0][1] * length_c, em->com[0][2] * length_c);
    std::fprintf(f, "\nTotal mass of the entire system \t= %f\n", em->mass[0]);
    {
        const double x = em->vel[0][0] * velocity_c, y = em->vel[0][1] * velocity_c, z = em->vel[0][2] * velocity_c;
        std::fprintf(f, "\nLinear velocity of the entire system system \t\t= (%f, %f, %f)", x, y, z);
        std::fprintf(f, "\nMagnitude of the linear velocity of the entire system \t= %f\n",
                     std::sqrt(x * x + y * y + z * z));
    }
    {
        const double x = em->angmom[0][0] * momentum_c, y = em->angmom[0][1] * momentum_c,
                     z = em->angmom[0][2] * momentum_c;
        std::fprintf(f, "\nAngular momentum of the entire system system \t\t= (%e, %e, %e)", x, y, z);
        std::fprintf(f, "\nMagnitude of the angular momentum of the entire system \t= %e\n",
                     std::sqrt(x * x + y * y + z * z));
    }
    std::fprintf(f, "\n\n*************************************************************************\n\n\n");
    return std::fclose(f) == 0 ? 0 : GI_E_IO;
}


int gi_spin_body(const gi_float4 *pos, gi_float4 *vel, int start, int stop, int stop_fe, double mass_fe,
                 double mass_si, double total_mass, const float *spin4) {
    if (!pos || !vel || !spin4 || start < 0 || stop < start || stop_fe < start || stop_fe > stop)
        return GI_E_INVALID_ARG;
    float sx = spin4[0], sy = spin4[1], sz = spin4[2];
    const float w = spin4[3];
    // The direction is normalised with a float norm (its operands are float).
    const double mag = std::sqrt(sx * sx + sy * sy + sz * sz);
    if (!(0.0000001 < mag)) return GI_E_INVALID_ARG;
    sx = (float)(sx / mag);
    sy = (float)(sy / mag);
    sz = (float)(sz / mag);
    double com[3];
    body_mean(pos, start, stop_fe, stop, mass_fe, mass_si, total_mass, com);
    for (int i = start; i < stop; ++i) {
        const double rx = pos[i].x - com[0], ry = pos[i].y - com[1], rz = pos[i].z - com[2];
        // distance from the spin axis, speed 2 pi w d along n = s x r / |s x r|
        const double r2 = rx * rx + ry * ry + rz * rz;
        const double along = sx * rx + sy * ry + sz * rz;
        const double d = std::sqrt(r2 - along * along);
        const double speed = w * 2.0 * kPi * d;
        double nx = (sy * rz - sz * ry);
        double ny = -(sx * rz - sz * rx);
        double nz = (sx * ry - sy * rx);
        const double nm = std::sqrt(nx * nx + ny * ny + nz * nz);
        if (nm != 0.0) {
            nx /= nm;
            ny /= nm;
            nz /= nm;
            vel[i].x = (float)(vel[i].x + speed * nx);
            vel[i].y = (float)(vel[i].y + speed * ny);
            vel[i].z = (float)(vel[i].z + speed * nz);
        }
    }
    return 0;
}

int gi_zero_out_drift(gi_float4 *pos, gi_float4 *vel, const gi_units *u) {
    if (!pos || !vel || !u) return GI_E_INVALID_ARG;
    const int b1 = u->nfe1, b2 = u->nfe1 + u->nsi1, b3 = b2 + u->nfe2, n = b3 + u->nsi2;
    const int ranges[2][3] = {{0, b1, b2}, {b2, b3, n}};
    const double mass[2] = {u->mass_of_body1, u->mass_of_body2};
    for (int b = 0; b < 2; ++b) {
        double com[3], v[3];
        body_mean(pos, ranges[b][0], ranges[b][1], ranges[b][2], u->mass_fe, u->mass_si, mass[b], com);
        body_mean(vel, ranges[b][0], ranges[b][1], ranges[b][2], u->mass_fe, u->mass_si, mass[b], v);
        for (int i = ranges[b][0]; i < ranges[b][2]; ++i) {
            pos[i].x = (float)(pos[i].x - com[0]);
            pos[i].y = (float)(pos[i].y - com[1]);
            pos[i].z = (float)(pos[i].z - com[2]);
            vel[i].x = (float)(vel[i].x - v[0]);
            vel[i].y = (float)(vel[i].y - v[1]);
            vel[i].z = (float)(vel[i].z - v[2]);
        }
    }
    return 0;
}

int gi_make_raw_bodies(const gi_units *u, unsigned long long seed, double jitter, gi_float4 *pos, gi_float4 *vel) {
    if (!u || !pos || !vel) return GI_E_INVALID_ARG;
    const int n1 = u->nfe1 + u->nsi1, n2 = u->nfe2 + u->nsi2;
    if (n1 <= 0 || n2 < 0) return GI_E_INVALID_ARG;
    // The reference draws the raw bodies with srand(time); here a seeded
    // jittered lattice with an iron core, both bodies centred at the origin,
    // in the generation layout [Fe1, Si1, Fe2, Si2] with its tags and masses
    // (generateTargetImpactorBasicFunctions.h:283-398).
    SplitMix64 rng{seed};
    const Body bodies[2] = {make_body_count(n1, u->nfe1, jitter, rng), make_body_count(n2, u->nfe2, jitter, rng)};
    int k = 0;
    for (int b = 0; b < 2; ++b) {
        const Body &bd = bodies[b];
        const int nb = b == 0 ? n1 : n2;
        for (int q = 0; q < nb; ++q, ++k) {
            const bool fe = q < bd.nfe;
            pos[k] = gi_float4{(float)bd.x[q], (float)bd.y[q], (float)bd.z[q], (float)(2 * b + (fe ? 0 : 1))};
            vel[k] = gi_float4{0.0f, 0.0f, 0.0f, (float)(fe ? u->mass_fe : u->mass_si)};
        }
    }
    return 0;
}

int gi_generate_bodies(const gi_generator_setup *setup, unsigned long long seed, double jitter, gi_float4 *pos,
                       gi_float4 *vel, gi_units *units, gi_generate_info *info, const char *out_dir) {
    if (!setup || !pos || !vel) return GI_E_INVALID_ARG;
    gi_units u;
    int rc = gi_units_compute(setup, &u);
    if (rc) return rc;
    const int n = setup->total_number_of_elements;
    const int n1 = u.nfe1 + u.nsi1, n2 = u.nfe2 + u.nsi2;
    if (n <= 0 || n1 + n2 != n) return GI_E_INVALID_ARG;
    gi_force_separate_constants fc;
    gi_move_separate_constants mc;
    gi_separate_constants(&u, setup->dt, &fc, &mc);
    if ((rc = gi_make_raw_bodies(&u, seed, jitter, pos, vel))) return rc;

    // Phase lengths with the reference's bookkeeping: after step k the time is
    // the double sum of k Dt's; the step moves with damping while that time is
    // below DampTime, rests until DampTime + DampRestTime, spins once, rests,
    // and the run ends after the step whose time exceeds the total
    // (generateTargetImpactor.cu:671-727).
    long long n_damp = 0, n_rest = 0, n_spin = 0;
    double t = 0.0;
    const double t_end = u.damp_time + u.damp_rest_time + u.spin_rest_time;
    for (;;) {
        t += setup->dt;
        if (t < u.damp_time)
            ++n_damp;
        else if (t < u.damp_time + u.damp_rest_time)
            ++n_rest;
        else
            ++n_spin;
        if (t_end < t) break;
        if (n_damp + n_rest + n_spin > (1LL << 31)) return GI_E_TOO_LARGE;
    }

    DeviceState d;
    if ((rc = d.alloc(n))) return rc;
    const size_t bytes = (size_t)n * 16;
    if ((rc = gi_copy_h2d(d.pos, pos, bytes)) || (rc = gi_copy_h2d(d.vel, vel, bytes))) return rc;
    gi_float4 *P = (gi_float4 *)d.pos, *V = (gi_float4 *)d.vel, *F = (gi_float4 *)d.force;
    if ((rc = gi_separate_steps(P, V, F, n, &fc, &mc, setup->damp_rate_body1, setup->damp_rate_body2, (int)n_damp,
                                nullptr)))
        return rc;
    if ((rc = gi_separate_steps(P, V, F, n, &fc, &mc, 0.0f, 0.0f, (int)n_rest, nullptr))) return rc;
    if (n_spin > 0) {
        // The spin step: force on the unspun state, then spin both bodies
        // (on the host, as spinBody), then move (generateTargetImpactor.cu:698-711).
        if ((rc = gi_force_separate(P, V, F, n, &fc, nullptr))) return rc;
        if ((rc = gi_copy_d2h(pos, P, bytes)) || (rc = gi_copy_d2h(vel, V, bytes))) return rc;
        const float s1[4] = {setup->initial_spin1[0], setup->initial_spin1[1], setup->initial_spin1[2], u.spin1_w};
        const float s2[4] = {setup->initial_spin2[0], setup->initial_spin2[1], setup->initial_spin2[2], u.spin2_w};
        if ((rc = gi_spin_body(pos, vel, 0, n1, u.nfe1, u.mass_fe, u.mass_si, u.mass_of_body1, s1))) return rc;
        if ((rc = gi_spin_body(pos, vel, n1, n, n1 + u.nfe2, u.mass_fe, u.mass_si, u.mass_of_body2, s2))) return rc;
        if ((rc = gi_copy_h2d(V, vel, bytes))) return rc;
        if ((rc = gi_move_damped_separate(P, V, F, n, &mc, 0.0f, 0.0f, nullptr))) return rc;
        if ((rc = gi_separate_steps(P, V, F, n, &fc, &mc, 0.0f, 0.0f, (int)(n_spin - 1), nullptr))) return rc;
    }
    if ((rc = gi_device_sync())) return rc;
    if ((rc = gi_copy_d2h(pos, P, bytes)) || (rc = gi_copy_d2h(vel, V, bytes))) return rc;
    if ((rc = gi_zero_out_drift(pos, vel, &u))) return rc;

    if (out_dir) {
        // recordCarryForwardParameters / recordTargetImpactorInitialPosVel
        // (generateTargetImpactorBasicFunctions.h:688-722).
        if ((rc = write_target_impactor_stats(join(out_dir, "targetImpactorStats"), *setup, u, pos, vel))) return rc;
        if ((rc = gi_parameters_write(join(out_dir, "targetImpactorParameters").c_str(), &u, n))) return rc;
        FILE *f = std::fopen(join(out_dir, "targetImpactorInitialPosVel").c_str(), "wb");
        if (!f) return GI_E_IO;
        rc = std::fwrite(&n, sizeof(int), 1, f) == 1 ? write_pos_vel(f, pos, vel, n) : GI_E_IO;
        if (std::fclose(f) != 0 && !rc) rc = GI_E_IO;
        if (rc) return rc;
    }
    if (units) *units = u;
    if (info) {
        info->n = n;
        info->damp_steps = n_damp;
        info->rest_steps = n_rest;
        info->spin_rest_steps = n_spin;
        info->run_time = t;
    }
    return 0;
}

int gi_impact_initialize(const char *in_dir, const char *setup_path, const char *out_dir) {
    if (!in_dir || !setup_path || !out_dir) return GI_E_INVALID_ARG;
    gi_units u;
    int n = 0;
    int rc = gi_parameters_read(join(in_dir, "targetImpactorParameters").c_str(), &u, &n);
    if (rc) return rc;
    if (n <= 0 || u.nfe1 + u.nsi1 + u.nfe2 + u.nsi2 != n) return GI_E_PARSE;
    // Body offsets (km) and velocities (km/s), positional
    // (impactInitialization.cu:188-217), stored as float like the reference's
    // float4 globals and converted to simulation units (:235-249).
    float ip1[3], ip2[3], iv1[3], iv2[3];
    {
        PositionalReader r(setup_path);
        if (!r.ok()) return GI_E_IO;
        bool ok = true;
        for (float *a : {ip1, ip2, iv1, iv2})
            for (int k = 0; k < 3; ++k) ok = ok && r.next(a[k]);
        if (!ok) return GI_E_PARSE;
    }
    for (int k = 0; k < 3; ++k) {
        ip1[k] = (float)(ip1[k] / u.unit_length);
        ip2[k] = (float)(ip2[k] / u.unit_length);
        iv1[k] = (float)(iv1[k] * (u.unit_time / u.unit_length));
        iv2[k] = (float)(iv2[k] * (u.unit_time / u.unit_length));
    }
    std::vector<gi_float4> tp(n), tv(n), pos(n), vel(n);
    {
        FILE *f = std::fopen(join(in_dir, "targetImpactorInitialPosVel").c_str(), "rb");
        if (!f) return GI_E_IO;
        int lead = 0;
        const bool ok = std::fread(&lead, sizeof(int), 1, f) == 1 &&
                        std::fread(tp.data(), sizeof(gi_float4), n, f) == (size_t)n &&
                        std::fread(tv.data(), sizeof(gi_float4), n, f) == (size_t)n;
        std::fclose(f);
        if (!ok) return GI_E_IO;
    }
    const int n1 = u.nfe1 + u.nsi1;
    for (int i = 0; i < n; ++i) {
        const float *dp = i < n1 ? ip1 : ip2;
        const float *dv = i < n1 ? iv1 : iv2;
        tp[i].x += dp[0];
        tp[i].y += dp[1];
        tp[i].z += dp[2];
        tv[i].x += dv[0];
        tv[i].y += dv[1];
        tv[i].z += dv[2];
    }
    // [Fe1, Si1, Fe2, Si2] -> [Fe1, Fe2, Si1, Si2] (impactInitialization.cu:329-371).
    const int seg[4][2] = {{0, u.nfe1}, {n1, n1 + u.nfe2}, {u.nfe1, n1}, {n1 + u.nfe2, n}};
    int k = 0;
    for (auto &s : seg)
        for (int i = s[0]; i < s[1]; ++i, ++k) {
            pos[k] = tp[i];
            vel[k] = tv[i];
        }
    {
        FILE *f = std::fopen(join(out_dir, "impactPosVel").c_str(), "wb");
        if (!f) return GI_E_IO;
        const float time0 = 0.0f;
        rc = std::fwrite(&time0, sizeof(float), 1, f) == 1 ? write_pos_vel(f, pos.data(), vel.data(), n) : GI_E_IO;
        if (std::fclose(f) != 0 && !rc) rc = GI_E_IO;
        if (rc) return rc;
    }
    // Centre of mass and mean velocity of the whole system in double
    // (impactInitialization.cu:395-446), kept as float for the recentring.
    const int nfe = u.nfe1 + u.nfe2;
    double mt = 0.0, c[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i) mt += i < nfe ? u.mass_fe : u.mass_si;
    for (int i = 0; i < n; ++i) {
        const double m = i < nfe ? u.mass_fe : u.mass_si;
        c[0] += pos[i].x * m;
        c[1] += pos[i].y * m;
        c[2] += pos[i].z * m;
    }
    for (int i = 0; i < n; ++i) {
        const double m = i < nfe ? u.mass_fe : u.mass_si;
        c[3] += vel[i].x * m;
        c[4] += vel[i].y * m;
        c[5] += vel[i].z * m;
    }
    for (int q = 0; q < 6; ++q) c[q] /= mt;
    // Angular momentum about the centre of mass, relative to the mean
    // velocity, in double (impactInitialization.cu:448-475).
    double L[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < n; ++i) {
        const double m = i < nfe ? u.mass_fe : u.mass_si;
        const double rx = pos[i].x - c[0], ry = pos[i].y - c[1], rz = pos[i].z - c[2];
        const double vx = vel[i].x - c[3], vy = vel[i].y - c[4], vz = vel[i].z - c[5];
        L[0] += (ry * vz - rz * vy) * m;
        L[1] += -(rx * vz - rz * vx) * m;
        L[2] += (rx * vy - ry * vx) * m;
    }
    if ((rc = write_initial_stats(join(out_dir, "initialStatsFile"), u, mt, c, c + 3, L))) return rc;
    float cf[6];
    for (int q = 0; q < 6; ++q) cf[q] = (float)c[q];
    return write_centering(join(out_dir, "viewCenteringInfo"), cf);
}

int gi_impact_run(const char *param_dir, const char *setup_path, const char *run_dir, gi_impact_run_info *info) {
    if (!param_dir || !setup_path || !run_dir) return GI_E_INVALID_ARG;
    gi_units u;
    int n = 0;
    int rc = gi_parameters_read(join(param_dir, "targetImpactorParameters").c_str(), &u, &n);
    if (rc) return rc;
    if (n <= 0 || u.nfe1 + u.nsi1 + u.nfe2 + u.nsi2 != n) return GI_E_PARSE;
    // setupImpactAddTime, positional (impactAddTime.cu:60-125).
    double additional_time = 0.0, dt = 0.0;
    int draw_rate = 0, record_rate = 0, recenter_rate = 0;
    {
        PositionalReader r(setup_path);
        if (!r.ok()) return GI_E_IO;
        if (!(r.next(additional_time) && r.next(dt) && r.next(draw_rate) && r.next(record_rate) &&
              r.next(recenter_rate)))
            return GI_E_PARSE;
    }
    // A rate of zero or less never fires: the reference's counters start at 1
    // after the first step and never equal it (impactAddTime.cu:402-404,420,433).
    if (!(dt > 0.0)) return GI_E_INVALID_ARG;
    additional_time *= 24.0 * 60.0 * 60.0;
    additional_time /= u.unit_time;
    float centering[6];
    {
        PositionalReader r(join(run_dir, "viewCenteringInfo").c_str());
        if (!r.ok()) return GI_E_IO;
        for (float &v : centering)
            if (!r.next(v)) return GI_E_PARSE;
    }
    gi_force_collision_constants fc;
    gi_move_collision_constants mc;
    gi_impact_constants(&u, dt, &fc, &mc);

    // Resume from the last frame (impactAddTime.cu:205-243).
    std::vector<gi_float4> pos(n), vel(n);
    const std::string frames = join(run_dir, "impactPosVel");
    FILE *ff = std::fopen(frames.c_str(), "rb+");
    if (!ff) return GI_E_IO;
    const long frame_bytes = (long)(sizeof(float) + 2 * (size_t)n * sizeof(gi_float4));
    float time0 = 0.0f;
    if (std::fseek(ff, -frame_bytes, SEEK_END) != 0 || std::fread(&time0, sizeof(float), 1, ff) != 1 ||
        std::fread(pos.data(), sizeof(gi_float4), n, ff) != (size_t)n ||
        std::fread(vel.data(), sizeof(gi_float4), n, ff) != (size_t)n) {
        std::fclose(ff);
        return GI_E_IO;
    }
    double run_time = (double)time0;
    const double total_run_time = run_time + additional_time;

    // Step count: the loop ends after the first step whose time exceeds the
    // total (impactAddTime.cu:405,457).
    long long n_steps = 0;
    {
        double t = run_time;
        do {
            t += dt;
            ++n_steps;
        } while (!(total_run_time < t));
    }

    DeviceState d;
    if ((rc = d.alloc(n))) {
        std::fclose(ff);
        return rc;
    }
    const size_t bytes = (size_t)n * 16;
    gi_float4 *P = (gi_float4 *)d.pos, *V = (gi_float4 *)d.vel, *F = (gi_float4 *)d.force;
    if ((rc = gi_copy_h2d(P, pos.data(), bytes)) || (rc = gi_copy_h2d(V, vel.data(), bytes))) {
        std::fclose(ff);
        return rc;
    }
    const int nfe = u.nfe1 + u.nfe2;
    const long long never = std::numeric_limits<long long>::max();
    int frames_written = 0;
    long long done = 0;
    const double start_time = run_time;
    while (done < n_steps) {
        // run to the next record / recentre event or the end
        const long long next_record = record_rate > 0 ? (done / record_rate + 1) * record_rate : never;
        const long long next_recenter = recenter_rate > 0 ? (done / recenter_rate + 1) * recenter_rate : never;
        const long long target = std::min(std::min(next_record, next_recenter), n_steps);
        const int k = (int)(target - done);
        if ((rc = gi_impact_steps(P, V, F, n, &fc, &mc, k, 0, nullptr, nullptr, nullptr))) break;
        for (int q = 0; q < k; ++q) run_time += dt;
        done = target;
        if (record_rate > 0 && done % record_rate == 0) {
            if ((rc = gi_device_sync())) break;
            if ((rc = gi_copy_d2h(pos.data(), P, bytes)) || (rc = gi_copy_d2h(vel.data(), V, bytes))) break;
            const float tf = (float)run_time;
            if (std::fseek(ff, 0, SEEK_END) != 0 || std::fwrite(&tf, sizeof(float), 1, ff) != 1 ||
                write_pos_vel(ff, pos.data(), vel.data(), n) != 0) {
                rc = GI_E_IO;
                break;
            }
            ++frames_written;
        }
        if (recenter_rate > 0 && done % recenter_rate == 0) {
            // the reference's float recenter toward the viewCenteringInfo
            // floats, with the %f-read element masses
            if ((rc = gi_recenter_float(P, V, n, nfe, u.mass_fe, u.mass_si, centering, centering + 3, nullptr)))
                break;
        }
    }
    if (!rc) rc = gi_device_sync();
    if (std::fclose(ff) != 0 && !rc) rc = GI_E_IO;
    if (rc) return rc;
    if ((rc = write_centering(join(run_dir, "viewCenteringInfo"), centering))) return rc;
    // The final statistics of a finished run (impactAddTime.cu:457-465): the
    // Earth / Moon finder on the device state, then runStatsFile.
    gi_earth_moon em;
    if ((rc = gi_find_earth_moon(P, V, n, u.nfe1, u.nfe2, u.nsi1, u.mass_fe, u.mass_si, (float)(u.diameter * 1.5),
                                 nullptr, &em, nullptr)))
        return rc;
    if ((rc = gi_write_run_stats(join(run_dir, "runStatsFile").c_str(), &em, &u, run_time))) return rc;
    if (info) {
        info->n = n;
        info->steps = n_steps;
        info->frames_written = frames_written;
        info->start_time = start_time;
        info->end_time = run_time;
    }
    return 0;
}

}  // extern "C"
