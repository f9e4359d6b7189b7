// NuOsc.h -- host C++ class with the interface of the reference's NuOsc
// (nuosc_class.h:99-191), backed by libnuosc_hip.so through the C ABI of
// include/nuosc_hip.h.  The state lives on the GPU; host code sees copies.
//
// Kept: constructor arguments, set_mu / set_pmo / set_renorm, get_lpts /
// get_nv, fillInitValue, step_rk4, step_srk3, calRHS, analysis (stdout line +
// analysis.dat), renormalize, addSnapShotAtV / addSnapShotAtXV / checkSnapShot
// and the public scalars phy_time, dt, dx, dz, ds_L, CFL, ko, pmo, mu, renorm,
// n_nue0, theta/ct/st.
// Changed: px, pz (MPI decomposition) must be 1; snapshots name their
// quantity ("P3" or a field name) instead of holding a raw host pointer,
// and read the device state at the time of the dump.
// Added: precision, velocity-grid weights, boundary (periodic or open z
// with an inflowing ipt-0 pulse centred at zc), the discretisation (the
// reference's KO_ORD_3 / ADVEC_* / N_FLAVOR_3 build options: ko_order, adv,
// nflavor) and, beside analysis.dat
// (whose columns stay the reference's), growth.dat with |rho_ex| per
// analysis call for linear growth rates.
#pragma once

#include <cmath>
#include <cstdint>
#include <fstream>
#include <list>
#include <string>
#include <vector>

#include "../../include/nuosc_hip.h"

struct NuOscGrid {
    int nx = 0, nz = 0, nv = 0, gx = 2, gz = 2;
    double x0 = 0, x1 = 0, z0 = 0, z1 = 0, dx = 0, dz = 0;
    std::vector<double> X, Z, vx, vy, vz, vw;
    long long lpts = 0;
    int get_nv() const { return nv; }
    long long get_lpts() const { return lpts; }
    // CartGrid.h:91
    long long idx(int i, int j, int v) const { return ((long long)(i + gx) * (nz + 2 * gz) + j + gz) * nv + v; }
};

struct NuOscSnapShot {
    std::vector<std::string> vars;   // "P3", or field names ee, mm, emr, ... (3 flavours: also tt, mtr, ...)
    std::string fntpl;
    int every = 1;
    std::vector<int> x_slices, v_slices;
};

class NuOsc {
public:
    int myrank = 0;
    double phy_time = 0, dt = 0, dx = 0, dz = 0;
    NuOscGrid grid;
    double ds_L = 0;
    double n_nue0[2] = {0, 0};
    double CFL = 0, ko = 0;
    const double theta = 37 * M_PI / 180.;
    const double ct = std::cos(2 * theta);
    const double st = std::sin(2 * theta);
    double pmo = 0.1;
    double mu = 1.0;
    bool renorm = false;
    int nvar = NUOSC_NVAR2;   // fields: 8 (2 flavours) or 18 (3 flavours, N_FLAVOR_3)

    NuOsc(int px, int pz, int nv, int nphi, int gx, int gz, double x0, double x1, double z0, double z1, double dx,
          double dz, double CFL, double ko, int precision = NUOSC_PREC_FP64, int vgrid = NUOSC_VGRID_VORONOI,
          bool use_graph = true, int device = 0, int bc = NUOSC_BC_PERIODIC, double zc = 0.0, int ko_order = 3,
          int adv = NUOSC_ADV_CFD4, int nflavor = 2);
    ~NuOsc();
    NuOsc(const NuOsc&) = delete;
    NuOsc& operator=(const NuOsc&) = delete;

    void set_mu(double mu_);
    void set_pmo(double pmo_);
    void set_renorm(bool renorm_);
    long long get_lpts() const { return grid.get_lpts(); }
    int get_nv() const { return grid.get_nv(); }

    void fillInitValue(int ipt, double alpha, double eps0, double sigma, double lnue, double lnueb, double lnuex,
                       double lnuebx, uint64_t seed = 0);
    void step_rk4();
    void step_rk4(int nsteps);
    void step_srk3();
    void step_srk3(int nsteps);
    void synchronize();
    // out/in: nvar fields in the interior layout [nx][nz][nv]
    void calRHS(std::vector<double>& out, const std::vector<double>& in);
    // prints the reference's stdout line and appends to analysis.dat;
    // returns the full diagnostics
    nuosc_analysis_t analysis();
    void renormalize();

    void addSnapShotAtV(const std::vector<std::string>& vars, const char* fntpl, int dumpstep, std::vector<int> vidx);
    void addSnapShotAtXV(const std::vector<std::string>& vars, const char* fntpl, int dumpstep,
                         std::vector<int> xidx, std::vector<int> vidx);
    void checkSnapShot(int t = 0);

    // host copy of one field, interior layout
    std::vector<double> field(int f);
    nuosc_sim* handle() { return sim_; }

private:
    nuosc_sim* sim_ = nullptr;
    std::ofstream anafile_, growthfile_;
    std::list<NuOscSnapShot> snapshots_;
    bool first_ana_line_ = true;
    std::vector<double> G0_, G0b_;
    void check(int rc, const char* what);
    void write_meta(const NuOscSnapShot& ss, bool x_then_z);
};
