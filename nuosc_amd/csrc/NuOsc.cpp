// NuOsc.cpp -- the reference-compatible host class over libnuosc_hip.so.
#include "NuOsc.h"

#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <stdexcept>

namespace {
// field indices of include/nuosc_hip.h (2 flavours: the first 8)
const char* kFieldNames[NUOSC_NVAR3] = {"ee",  "mm",  "emr", "emi",  "bee",  "bmm",  "bemr", "bemi", "tt",
                                        "mtr", "mti", "ter", "tei",  "btt",  "bmtr", "bmti", "bter", "btei"};

int field_index(const std::string& name, int nvar) {
    for (int f = 0; f < nvar; ++f)
        if (name == kFieldNames[f]) return f;
    if (name == "xx") return 1;  // the reference also calls mm "xx"
    return -1;
}
}  // namespace

void NuOsc::check(int rc, const char* what) {
    if (rc != NUOSC_OK) {
        std::string msg = std::string(what) + ": " + nuosc_last_error(sim_);
        throw std::runtime_error(msg);
    }
}

NuOsc::NuOsc(int px, int pz, int nv, int nphi, int gx, int gz, double x0, double x1, double z0, double z1,
             double dx_, double dz_, double CFL_, double ko_, int precision, int vgrid, bool use_graph, int device,
             int bc, double zc, int ko_order, int adv, int nflavor)
    : dx(dx_), dz(dz_), CFL(CFL_), ko(ko_) {
    (void)nphi;  // polar velocity grid only
    if (px != 1 || pz != 1)
        throw std::runtime_error("one domain per GPU: --np must be 1 1");
    nuosc_config c{};
    c.nn = nv;
    c.gx = gx;
    c.gz = gz;
    c.x0 = x0;
    c.x1 = x1;
    c.z0 = z0;
    c.z1 = z1;
    c.dx = dx_;
    c.dz = dz_;
    c.cfl = CFL_;
    c.ko = ko_;
    c.mu = mu;
    c.pmo = pmo;
    c.renorm = 0;
    c.vgrid = vgrid;
    c.precision = precision;
    c.bc = bc;
    c.zc = zc;
    c.ko_order = ko_order;
    c.adv = adv;
    c.nflavor = nflavor;
    nvar = nflavor == 3 ? NUOSC_NVAR3 : NUOSC_NVAR2;
    c.device = device;
    c.use_graph = use_graph ? 1 : 0;
    int rc = nuosc_create(&c, &sim_);
    if (rc != NUOSC_OK) throw std::runtime_error(std::string("nuosc_create: ") + nuosc_last_error(nullptr));

    int nx_ = 0, nz_ = 0, nv_ = 0;
    long long lpts = 0;
    nuosc_get_dims(sim_, &nx_, &nz_, &nv_, &lpts);
    grid.nx = nx_;
    grid.nz = nz_;
    grid.nv = nv_;
    grid.gx = gx;
    grid.gz = gz;
    grid.x0 = x0;
    grid.x1 = x1;
    grid.z0 = z0;
    grid.z1 = z1;
    grid.dx = dx_;
    grid.dz = dz_;
    grid.lpts = lpts;
    grid.X.resize(nx_);
    grid.Z.resize(nz_);
    nuosc_get_coords(sim_, grid.X.data(), grid.Z.data());
    grid.vx.resize(nv_);
    grid.vy.resize(nv_);
    grid.vz.resize(nv_);
    grid.vw.resize(nv_);
    nuosc_get_vgrid(sim_, grid.vx.data(), grid.vy.data(), grid.vz.data(), grid.vw.data());
    nuosc_get_scalars(sim_, &dt, &ds_L, &phy_time, n_nue0);

    long long bytes = 0;
    nuosc_get_device_bytes(sim_, &bytes);
    // nuosc_class.cpp:473-510
    std::printf("\nNuOsc2D on 1 GPU (%s).\n", precision == NUOSC_PREC_FP64 ? "FP64" : "FP32 + FP64 moments");
    std::printf("   Domain:  v: nv = %5d  nphi = %5d  on S2.\n", grid.nv, 0);
    std::printf("            x:( %12f %12f )  dx = %g\n", x0, x1, dx);
    std::printf("            z:( %12f %12f )  dz = %g\n", z0, z1, dz);
    std::printf("   Fieldvar: %d -- device memory %.2f GB\n", nvar, bytes / 1024. / 1024. / 1024.);
    std::printf("   dt = %g     CFL = %g\n", dt, CFL);
    if (bc == NUOSC_BC_OPEN_Z)
        std::printf("   Boundary: x periodic, z open (ipt-0 inflow pulse centred at z = %g)\n", zc);
    else
        std::printf("   Boundary: Periodic\n");
    std::printf("   Velocity grid: Icosahedron sphere grid%s\n", vgrid == NUOSC_VGRID_EQUAL ? " (equal weights)" : "");
    // nuosc_class.cpp:492-503
    if (ko_order == 5) std::printf("   Dissipation: 5-th order KO w/ eps = %g\n", ko);
    else std::printf("   Dissipation: 3-th order KO w/ eps = %g\n", ko);
    static const char* const kAdvName[] = {"ON (Center-FD)", "ON (Center-FD, 2nd order)", "ON (2nd-order upwind)",
                                           "ON (4th-order lopsided)", "ON (MUSCL, superbee)", "OFF"};
    std::printf("   Advection: %s\n", kAdvName[adv >= 0 && adv <= NUOSC_ADV_OFF ? adv : 0]);
    std::printf("   Vacuum term: pmo= %g theta= %g\n", pmo, theta);

    // nuosc_class.cpp:536-540
    anafile_.open("analysis.dat", std::ofstream::out | std::ofstream::trunc);
    if (!anafile_) std::cout << "*** Open fails: " << "./analysis.dat" << std::endl;
    anafile_ << "### [ phy_time,   1:maxrelP,    2:surv, survb,    4:avgP, avgPb,      6:aM0    7:Lem   8:ELNe]"
             << std::endl;
    growthfile_.open("growth.dat", std::ofstream::out | std::ofstream::trunc);
    growthfile_ << "### [ phy_time,   1:|rho_ex| = sqrt(ds_L * sum w (emr^2 + emi^2)) ]" << std::endl;
}

NuOsc::~NuOsc() {
    anafile_.close();
    growthfile_.close();
    nuosc_destroy(sim_);
}

void NuOsc::set_mu(double mu_) {
    mu = mu_;
    check(nuosc_set_mu(sim_, mu), "set_mu");
    std::printf("   Setting mu = %f\n", mu);
}

void NuOsc::set_pmo(double pmo_) {
    pmo = pmo_;
    check(nuosc_set_pmo(sim_, pmo), "set_pmo");
    std::printf("   Setting pmo = %f\n", pmo);
}

void NuOsc::set_renorm(bool renorm_) {
    renorm = renorm_;
    check(nuosc_set_renorm(sim_, renorm ? 1 : 0), "set_renorm");
    std::printf("   Setting renorm = %d\n", renorm);
}

void NuOsc::fillInitValue(int ipt, double alpha, double eps0, double sigma, double lnue, double lnueb,
                          double lnuex, double lnuebx, uint64_t seed) {
    // nuosc_init.cpp:72
    std::printf("   Init data: [%s] alpha= %f eps= %g sigma= %g lnu_z= %g %g lnu_x= %g %g\n",
                ipt == 0 ? "Point-like pertur" : (ipt == 5 ? "Seeded complex pertur" : "Random pertur"), alpha,
                eps0, sigma, lnue, lnueb, lnuex, lnuebx);
    check(nuosc_fill_init(sim_, ipt, alpha, eps0, sigma, lnue, lnueb, lnuex, lnuebx, seed), "fillInitValue");
    nuosc_get_scalars(sim_, &dt, &ds_L, &phy_time, n_nue0);
    G0_.assign(grid.nv, 0.0);
    G0b_.assign(grid.nv, 0.0);
    nuosc_get_g0(sim_, G0_.data(), G0b_.data());
    std::printf("      init number density of nu_e: %g %g\n", n_nue0[0], n_nue0[1]);
}

void NuOsc::step_rk4() { step_rk4(1); }

void NuOsc::step_rk4(int nsteps) {
    check(nuosc_step_rk4(sim_, nsteps), "step_rk4");
    nuosc_get_scalars(sim_, nullptr, nullptr, &phy_time, nullptr);
}

void NuOsc::step_srk3() { step_srk3(1); }

void NuOsc::step_srk3(int nsteps) {
    check(nuosc_step_srk3(sim_, nsteps), "step_srk3");
    nuosc_get_scalars(sim_, nullptr, nullptr, &phy_time, nullptr);
}

void NuOsc::synchronize() { check(nuosc_synchronize(sim_), "synchronize"); }

void NuOsc::calRHS(std::vector<double>& out, const std::vector<double>& in) {
    const size_t n = (size_t)nvar * grid.nx * grid.nz * grid.nv;
    if (in.size() != n) throw std::runtime_error("calRHS: input size mismatch");
    out.assign(n, 0.0);
    check(nuosc_calrhs(sim_, in.data(), out.data()), "calRHS");
}

nuosc_analysis_t NuOsc::analysis() {
    nuosc_analysis_t a{};
    int rc = nuosc_analysis(sim_, &a);
    if (rc != NUOSC_OK && rc != NUOSC_EUNSTABLE) check(rc, "analysis");
    // nuosc_ana.cpp:119-129
    std::printf("T= %15f ", a.phy_time);
    std::printf(" |dP|max= %5.4e surb= %5.4e %5.4e conP= %5.4e %5.4e |M0|= %5.4e ELNe= %g Lex= %g\n", a.maxdP, a.surv,
                a.survb, a.avgP, a.avgPb, a.aM0, a.ELNe, a.Lex);
    anafile_ << a.phy_time << std::setprecision(13) << " " << a.maxdP << " " << a.surv << " " << a.survb << " "
             << a.avgP << " " << a.avgPb << " " << a.aM0 << " " << a.Lex << " " << a.ELNe << " " << std::endl;
    first_ana_line_ = false;
    growthfile_ << std::setprecision(13) << a.phy_time << " " << std::sqrt(a.rho_ex2) << std::endl;
    if (a.status != NUOSC_OK) {
        // the reference asserts maxdP < 10 (nuosc_ana.cpp:131); a NaN state also stops here
        std::fprintf(stderr, "analysis: maxdP = %g (non-finite points: %g) -- the run is unstable, stopping.\n",
                     a.maxdP, a.nonfinite);
        anafile_.close();
        std::exit(3);
    }
    return a;
}

void NuOsc::renormalize() { check(nuosc_renormalize(sim_), "renormalize"); }

std::vector<double> NuOsc::field(int f) {
    std::vector<double> out((size_t)grid.nx * grid.nz * grid.nv);
    check(nuosc_download_field(sim_, f, out.data()), "download_field");
    return out;
}

void NuOsc::write_meta(const NuOscSnapShot& ss, bool x_then_z) {
    // nuosc_snapshot.cpp:25-41 (AtXV) and 59-75 (AtV)
    char filename[256];
    const std::string tmp = ss.fntpl + ".meta";
    std::snprintf(filename, sizeof(filename), tmp.c_str(), 0);
    std::ofstream outfile(filename, std::ofstream::out | std::ofstream::trunc);
    if (!outfile) std::cout << "*** Open fails: " << filename << std::endl;
    const int sx = (int)ss.x_slices.size(), sv = (int)ss.v_slices.size();
    outfile << dt << " " << sx << " " << grid.nz << " " << sv << std::endl;
    outfile << grid.x0 << " " << grid.x1 << std::endl;
    outfile << grid.z0 << " " << grid.z1 << std::endl;
    if (x_then_z) {
        for (int i = 0; i < grid.nx; ++i) outfile << grid.X[i] << " ";
        outfile << std::endl;
        for (int i = 0; i < grid.nz; ++i) outfile << grid.Z[i] << " ";
        outfile << std::endl;
    } else {
        for (int i = 0; i < grid.nz; ++i) outfile << grid.Z[i] << " ";
        outfile << std::endl;
        for (int x : ss.x_slices) outfile << grid.X[x] << " ";
        outfile << std::endl;
    }
    for (int v : ss.v_slices) outfile << grid.vx[v] << " ";
    outfile << std::endl;
    for (int v : ss.v_slices) outfile << grid.vy[v] << " ";
    outfile << std::endl;
    for (int v : ss.v_slices) outfile << grid.vz[v] << " ";
    outfile << std::endl;
}

void NuOsc::addSnapShotAtXV(const std::vector<std::string>& vars, const char* fntpl, int dumpstep,
                            std::vector<int> xidx, std::vector<int> vidx) {
    NuOscSnapShot ss{vars, fntpl, dumpstep, xidx, vidx};
    for (int x : xidx)
        if (x < 0 || x >= grid.nx) throw std::runtime_error("snapshot x index out of range");
    for (int v : vidx)
        if (v < 0 || v >= grid.nv) throw std::runtime_error("snapshot v index out of range");
    std::printf("Add %d x %d x %d (XxZxV) snapshot every %d steps.\n", (int)xidx.size(), grid.nz, (int)vidx.size(),
                dumpstep);
    write_meta(ss, false);
    snapshots_.push_back(ss);
}

void NuOsc::addSnapShotAtV(const std::vector<std::string>& vars, const char* fntpl, int dumpstep,
                           std::vector<int> vidx) {
    std::vector<int> xidx(grid.nx);
    for (int i = 0; i < grid.nx; ++i) xidx[i] = i;
    NuOscSnapShot ss{vars, fntpl, dumpstep, xidx, vidx};
    for (int v : vidx)
        if (v < 0 || v >= grid.nv) throw std::runtime_error("snapshot v index out of range");
    std::printf("Add %d x %d x %d (XxZxV) snapshot every %d steps.\n", grid.nx, grid.nz, (int)vidx.size(), dumpstep);
    write_meta(ss, true);
    snapshots_.push_back(ss);
}

void NuOsc::checkSnapShot(int t) {
    // nuosc_snapshot.cpp:82-122: int32 step, f64 time, then per variable
    // f64[sx][nz][sv]
    for (const auto& ss : snapshots_) {
        const int sx = (int)ss.x_slices.size(), sv = (int)ss.v_slices.size();
        char filename[256];
        std::snprintf(filename, sizeof(filename), ss.fntpl.c_str(), t);
        std::ofstream outfile(filename, std::ofstream::out | std::ofstream::trunc | std::ofstream::binary);
        if (!outfile) std::cout << "*** Open fails: " << filename << std::endl;
        std::printf("\t\tWriting %d vars of size %d x %d x %d (XxZxV) into %s\n", (int)ss.vars.size(), sx, grid.nz, sv,
                    filename);
        const int32_t t32 = t;
        outfile.write(reinterpret_cast<const char*>(&t32), sizeof(int32_t));
        outfile.write(reinterpret_cast<const char*>(&phy_time), sizeof(double));
        std::vector<double> carr((size_t)sx * grid.nz * sv);
        for (const auto& var : ss.vars) {
            if (var == "P3") {
                check(nuosc_p3_slice(sim_, ss.x_slices.data(), sx, ss.v_slices.data(), sv, carr.data()), "P3 slice");
            } else {
                const int f = field_index(var, nvar);
                if (f < 0) throw std::runtime_error("unknown snapshot variable " + var);
                const std::vector<double> fld = field(f);
                for (int a = 0; a < sx; ++a)
                    for (int j = 0; j < grid.nz; ++j)
                        for (int b = 0; b < sv; ++b)
                            carr[((size_t)a * grid.nz + j) * sv + b] =
                                fld[((size_t)ss.x_slices[a] * grid.nz + j) * grid.nv + ss.v_slices[b]];
            }
            outfile.write(reinterpret_cast<const char*>(carr.data()), sizeof(double) * carr.size());
        }
    }
}
