// Command-line driver around the reference IroIro++ Wilson solve path.
//
// Test infrastructure only: this file is compiled together with the
// reference's own translation units (see oracle/ref_build.sh) into
// oracle/_ref/iroiro_ref.  It is used to write golden fixtures and to time
// the reference CPU path; nothing in the product links or runs it.
//
// Every operator call below goes through the reference classes unchanged:
//   Dirac_Wilson_EvenOdd<Dirac_Wilson>  lib/Dirac_ops/dirac_wilson_EvenOdd.hpp:12-92
//   Dirac_Wilson (full index)           lib/Dirac_ops/dirac_wilson.hpp:118-134
//   Fopr_DdagD                          lib/Fopr/fopr.h:146-153
//   Solver_CG                           lib/Solver/solver_CG.cpp:29-160
//   MultiShiftSolver_CG                 lib/Solver/multiShiftSolver_CG.cpp:10-185
//   Inverter_WilsonLike::invert         lib/Dirac_ops/eoUtils.cpp:43-80
//   AntiPeriodicBC<GaugeField>          lib/Dirac_ops/BoundaryConditions/antiPeriodicBC.hpp:25-37
//   Dirac_Wilson(_EvenOdd)::md_force    lib/Dirac_ops/dirac_wilson.cpp:41-92,
//                                       dirac_wilson_EvenOdd.hpp:94-116
//   FieldUtils::TracelessAntihermite    lib/Tools/fieldUtils.cpp:264-274
//   CCIO::BinaryReader / BinaryWriter   lib/IO/binary_reader.cpp, binary_writer.cpp
//   CCIO::NERSCWriter                   lib/IO/NERSC_writer.cpp
//   MesonCorrelator::calculate(_mom, _PSmom)
//                                       lib/Measurements/FermionicM/meson_correlator.hpp
//   (driven row by row as CCIO::ReadFromDisk / SaveOnDisk do on one
//   process, lib/IO/fields_io.hpp:47-138, 250-360; that header itself pulls
//   in the ILDG reader, which needs the LIME library)
//
// All binary files are raw little-endian float64 in the reference layouts
// (Format_F spinors, Format_G links, even/odd half fields in esec/osec order).
#include "include/commonPrms.hpp"
#include "include/common_fields.hpp"
#include "Communicator/communicator.hpp"
#include "Geometry/siteIndex.hpp"
#include "Geometry/siteIndex_EvenOdd.hpp"
#include "Dirac_ops/dirac_wilson.hpp"
#include "Dirac_ops/dirac_wilson_EvenOdd.hpp"
#include "Dirac_ops/eoUtils.hpp"
#include "Dirac_ops/BoundaryConditions/antiPeriodicBC.hpp"
#include "Fopr/fopr.h"
#include "Solver/solver_CG.hpp"
#include "Solver/multiShiftSolver_CG.hpp"
#include "Tools/fieldUtils.hpp"
#include "IO/binary_reader.hpp"
#include "IO/binary_writer.hpp"
#include "IO/NERSC_writer.hpp"
#include "Measurements/FermionicM/meson_correlator.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>
#include <sys/time.h>
#include <unistd.h>

namespace {

void die(const char* msg) {
  std::fprintf(stderr, "iroiro_ref: %s\n", msg);
  std::exit(2);
}

Field read_field(const std::string& path, size_t n) {
  Field f(n);
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) die(("cannot open " + path).c_str());
  size_t got = std::fread(f.getaddr(0), sizeof(double), n, fp);
  std::fclose(fp);
  if (got != n) die(("short read from " + path).c_str());
  return f;
}

void write_field(const std::string& path, const Field& f) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) die(("cannot create " + path).c_str());
  Field& g = const_cast<Field&>(f);
  std::fwrite(g.getaddr(0), sizeof(double), f.size(), fp);
  std::fclose(fp);
}

void write_vec(const std::string& path, const std::vector<double>& v) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) die(("cannot create " + path).c_str());
  std::fwrite(&v[0], sizeof(double), v.size(), fp);
  std::fclose(fp);
}

double now_ms() {
  timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec * 1000.0 + tv.tv_usec / 1000.0;
}

// Deterministic random numbers for timing runs (xorshift64), values in
// [-1, 1).
struct Filler {
  unsigned long long s;
  explicit Filler(unsigned long long seed) : s(seed) {}
  double next() {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
  }
};

// Random SU(3) link: Gram-Schmidt on two random complex rows, third row
// = conj(row0 x row1), so that the matrix is unitary with det 1.
void random_su3(Filler& rng, double* m) {
  double a[3][2], b[3][2];
  for (int i = 0; i < 3; ++i) { a[i][0] = rng.next(); a[i][1] = rng.next(); }
  for (int i = 0; i < 3; ++i) { b[i][0] = rng.next(); b[i][1] = rng.next(); }
  double na = 0; for (int i = 0; i < 3; ++i) na += a[i][0]*a[i][0] + a[i][1]*a[i][1];
  na = 1.0 / std::sqrt(na);
  for (int i = 0; i < 3; ++i) { a[i][0] *= na; a[i][1] *= na; }
  double pr = 0, pi = 0;  // <a,b>
  for (int i = 0; i < 3; ++i) {
    pr += a[i][0]*b[i][0] + a[i][1]*b[i][1];
    pi += a[i][0]*b[i][1] - a[i][1]*b[i][0];
  }
  for (int i = 0; i < 3; ++i) {
    b[i][0] -= pr*a[i][0] - pi*a[i][1];
    b[i][1] -= pr*a[i][1] + pi*a[i][0];
  }
  double nb = 0; for (int i = 0; i < 3; ++i) nb += b[i][0]*b[i][0] + b[i][1]*b[i][1];
  nb = 1.0 / std::sqrt(nb);
  for (int i = 0; i < 3; ++i) { b[i][0] *= nb; b[i][1] *= nb; }
  for (int j = 0; j < 3; ++j) {
    m[2*j] = a[j][0]; m[2*j+1] = a[j][1];
    m[6+2*j] = b[j][0]; m[6+2*j+1] = b[j][1];
  }
  for (int j = 0; j < 3; ++j) {
    int k = (j+1)%3, l = (j+2)%3;
    double cr = a[k][0]*b[l][0] - a[k][1]*b[l][1] - (a[l][0]*b[k][0] - a[l][1]*b[k][1]);
    double ci = a[k][0]*b[l][1] + a[k][1]*b[l][0] - (a[l][0]*b[k][1] + a[l][1]*b[k][0]);
    m[12+2*j] = cr; m[12+2*j+1] = -ci;
  }
}

void setup_lattice(char** argv) {
  Lattice latt;
  latt.stdinput = false;
  latt.Lx = std::atoi(argv[0]);
  latt.Ly = std::atoi(argv[1]);
  latt.Lz = std::atoi(argv[2]);
  latt.Lt = std::atoi(argv[3]);
  latt.NPEx = latt.NPEy = latt.NPEz = latt.NPEt = 1;
  CommonPrms::instance(latt);
  Communicator::instance();
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 7) {
    std::fprintf(stderr,
      "usage: iroiro_ref ops    Lx Ly Lz Lt mass gauge in_full outdir\n"
      "       iroiro_ref apbc   Lx Ly Lz Lt gauge_in gauge_out\n"
      "       iroiro_ref cg     Lx Ly Lz Lt mass gauge src_even prec maxiter outdir\n"
      "       iroiro_ref mshift Lx Ly Lz Lt mass gauge src_even prec maxiter outdir s0 s1 ...\n"
      "       iroiro_ref invert Lx Ly Lz Lt mass gauge src_full prec maxiter outdir\n"
      "       iroiro_ref mdforce Lx Ly Lz Lt mass gauge eta_even zeta_even eta_full zeta_full outdir\n"
      "       iroiro_ref iowrite Lx Ly Lz Lt Binary|NERSC gauge file\n"
      "       iroiro_ref ioread  Lx Ly Lz Lt Binary file gauge_out\n"
      "       iroiro_ref meson   Lx Ly Lz Lt q1 q2 px py pz outdir\n"
      "       iroiro_ref bench  Lx Ly Lz Lt mass nrep [cg_prec]\n"
      "       iroiro_ref benchprop Lx Ly Lz Lt mass gauge nrep cg_prec\n");
    return 2;
  }
  std::string cmd = argv[1];
  setup_lattice(argv + 2);
  const int V = CommonPrms::Nvol();
  const int Vh = V / 2;
  const size_t gsz = (size_t)18 * V * 4;
  const size_t fsz = (size_t)24 * V;
  const size_t hsz = (size_t)24 * Vh;
  SiteIndex_EvenOdd* eo = SiteIndex_EvenOdd::instance();

  if (cmd == "apbc") {
    GaugeField U;
    U.data = read_field(argv[6], gsz);
    AntiPeriodicBC<GaugeField>(TDIR).apply_bc(U);
    write_field(argv[7], U.data);
    return 0;
  }

  if (cmd == "iowrite" || cmd == "ioread") {
    // the row loop of SaveOnDisk / ReadFromDisk on one process: rows of Lx
    // sites in (t, z, y) order, each row one reader/writer call of
    // Nin*Lx*Nex doubles arranged by format(x, in, ex)
    const std::string id = argv[6];
    const int Lx = CommonPrms::Lx(), Ly = CommonPrms::Ly(), Lz = CommonPrms::Lz();
    const int Lt = CommonPrms::Lt(), Nin = 18, Nex = 4;
    std::vector<double> row((size_t)Nin * Lx * Nex);
    if (cmd == "iowrite") {
      Field u = read_field(argv[7], gsz);
      CCIO::GeneralWriter* w = 0;
      if (id == "Binary") w = new CCIO::BinaryWriter(V, Nin, Nex);
      else if (id == "NERSC") w = new CCIO::NERSCWriter(V, Nin, Nex, u);
      else die("unknown writer");
      FILE* fp = std::fopen(argv[8], "w");
      if (!fp) die("cannot create output");
      w->set_output(fp);
      w->header();
      for (int t = 0; t < Lt; ++t)
        for (int z = 0; z < Lz; ++z)
          for (int y = 0; y < Ly; ++y) {
            for (int x = 0; x < Lx; ++x)
              for (int ex = 0; ex < Nex; ++ex)
                for (int in = 0; in < Nin; ++in) {
                  const int site = x + Lx * (y + Ly * (z + Lz * t));
                  row[w->format(x, in, ex)] = u[in + Nin * (site + V * ex)];
                }
            w->write(&row[0], (unsigned)row.size());
          }
      std::fclose(fp);
      std::printf("IIREF_OK iowrite %s\n", id.c_str());
      return 0;
    }
    if (id != "Binary") die("unknown reader");
    CCIO::BinaryReader r(V, Nin, Nex);
    FILE* fp = std::fopen(argv[7], "r");
    if (!fp) die("cannot open input");
    r.set_sources(fp);
    r.header();
    Field u(gsz);
    for (int t = 0; t < Lt; ++t)
      for (int z = 0; z < Lz; ++z)
        for (int y = 0; y < Ly; ++y) {
          r.read(&row[0], (unsigned)row.size());
          for (int x = 0; x < Lx; ++x)
            for (int ex = 0; ex < Nex; ++ex)
              for (int in = 0; in < Nin; ++in) {
                const int site = x + Lx * (y + Ly * (z + Lz * t));
                u.set(in + Nin * (site + V * ex), row[r.format(x, in, ex)]);
              }
        }
    std::fclose(fp);
    write_field(argv[8], u);
    std::printf("IIREF_OK ioread %s\n", id.c_str());
    return 0;
  }

  if (cmd == "meson") {
    // q1, q2: 12 full fields each (columns c + 3 s), back to back
    prop_t q1(12), q2(12);
    Field a = read_field(argv[6], 12 * fsz), b = read_field(argv[7], 12 * fsz);
    for (int k = 0; k < 12; ++k) {
      q1[k] = Field(fsz);
      q2[k] = Field(fsz);
      for (size_t i = 0; i < fsz; ++i) {
        q1[k].set(i, a[k * fsz + i]);
        q2[k].set(i, b[k * fsz + i]);
      }
    }
    const double px = std::atof(argv[8]), py = std::atof(argv[9]), pz = std::atof(argv[10]);
    std::string out = argv[11];
    for (int type = 0; type <= PS_AV4; ++type) {
      char name[64];
      for (int dir = 0; dir < 4; ++dir) {
        MesonCorrelator mc((MesonType)type);
        std::vector<double> c = mc.calculate<Format::Format_F>(q1, q2, dir);
        std::snprintf(name, sizeof name, "/meson_%d_dir%d.bin", type, dir);
        write_vec(out + name, c);
      }
      MesonCorrelator mm((MesonType)type);
      std::vector<double> cm = mm.calculate_mom<Format::Format_F>(q1, q2, px, py, pz);
      std::snprintf(name, sizeof name, "/meson_%d_mom.bin", type);
      write_vec(out + name, cm);
      MesonCorrelator mp((MesonType)type);
      std::vector<double> cp = mp.calculate_PSmom<Format::Format_F>(q1, q2, px, py, pz);
      std::snprintf(name, sizeof name, "/meson_%d_psmom.bin", type);
      write_vec(out + name, cp);
    }
    std::printf("IIREF_OK meson\n");
    return 0;
  }

  if (cmd == "benchprop") {
    // The reference CPU path on a given gauge field (with its boundary
    // phase): one EO hop (Doe.mult), then the Schur solve of column 0 of
    // Qprop_EvenOdd's point source at the origin exactly as
    // Inverter_WilsonLike::invert calls it (eoUtils.cpp:51-60: Solver_CG on
    // Fopr_DdagD for Dhat^dag bhat, bhat = b_e since b_o = 0 for a source on
    // an even site), stopped at cg_prec (squared) so that a bounded sample
    // of iterations is timed.
    double mass = std::atof(argv[6]);
    Field u = read_field(argv[7], gsz);
    int nrep = std::atoi(argv[8]);
    double prec = std::atof(argv[9]);
    Dirac_Wilson_EvenOdd<Dirac_Wilson> D(mass, &u);
    Field be(hsz);
    be.set(0, 1.0);  // site 0 (even, esec index 0), spin 0, colour 0, real part
    Field w = D.mult_oe(be);  // warm-up
    double t0 = now_ms();
    for (int r = 0; r < nrep; ++r) w = D.mult_oe(be);
    double hop_ms = (now_ms() - t0) / nrep;
    Fopr_DdagD op(&D);
    double t2 = now_ms();
    w = op.mult(be);
    double ddagd_ms = now_ms() - t2;
    Field rhs = D.mult_dag(be);
    Solver_CG cg(prec, 100000, &op);
    Field x(hsz);
    SolverOutput so = cg.solve(x, rhs);
    std::printf("IIREF_BENCH {\"hop_ms\": %.6f, \"nrep\": %d, \"Vh\": %d, "
                "\"ddagd_ms\": %.6f, \"cg_ms\": %.6f, \"cg_iterations\": %d}\n",
                hop_ms, nrep, Vh, ddagd_ms, (double)so.timing, so.Iterations);
    return 0;
  }

  if (cmd == "bench") {
    // Times the reference EO hop (Doe.mult, one hop over V/2 sites) and a
    // complete Solver_CG run on a random SU(3) field with a loose stopping
    // precision, so that a bounded sample of CG iterations is measured.
    double mass = std::atof(argv[6]);
    int nrep = std::atoi(argv[7]);
    double prec = argc > 8 ? std::atof(argv[8]) : 0.0;
    Filler rng(0x9E3779B97F4A7C15ULL);
    Field u(gsz), fe(hsz);
    for (size_t l = 0; l < gsz / 18; ++l) random_su3(rng, u.getaddr(18 * l));
    for (size_t i = 0; i < hsz; ++i) fe.set(i, rng.next());
    Dirac_Wilson_EvenOdd<Dirac_Wilson> D(mass, &u);
    Field w = D.mult_oe(fe);  // warm-up
    double t0 = now_ms();
    for (int r = 0; r < nrep; ++r) w = D.mult_oe(fe);
    double hop_ms = (now_ms() - t0) / nrep;
    double ddagd_ms = -1.0, cg_ms = -1.0;
    int iters = -1;
    if (prec > 0.0) {
      Fopr_DdagD op(&D);
      double t2 = now_ms();
      w = op.mult(fe);
      ddagd_ms = now_ms() - t2;
      Solver_CG cg(prec, 100000, &op);
      Field x(hsz);
      SolverOutput so = cg.solve(x, fe);
      cg_ms = so.timing;
      iters = so.Iterations;
    }
    std::printf("IIREF_BENCH {\"hop_ms\": %.6f, \"nrep\": %d, \"Vh\": %d, "
                "\"ddagd_ms\": %.6f, \"cg_ms\": %.6f, \"cg_iterations\": %d}\n",
                hop_ms, nrep, Vh, ddagd_ms, cg_ms, iters);
    return 0;
  }

  double mass = std::atof(argv[6]);
  Field u = read_field(argv[7], gsz);
  Dirac_Wilson_EvenOdd<Dirac_Wilson> D(mass, &u);

  if (cmd == "ops") {
    std::string out = argv[9];
    Field ff = read_field(argv[8], fsz);
    Format::Format_F fmt(V);
    Field fe(ff[fmt.get_sub(eo->esec())]);
    Field fo(ff[fmt.get_sub(eo->osec())]);
    write_field(out + "/deo.bin", D.mult_eo(fo));
    write_field(out + "/doe.bin", D.mult_oe(fe));
    write_field(out + "/deo_dag.bin", D.mult_eo_dag(fe));
    write_field(out + "/doe_dag.bin", D.mult_oe_dag(fo));
    write_field(out + "/dhat.bin", D.mult(fe));
    write_field(out + "/dhat_dag.bin", D.mult_dag(fe));
    write_field(out + "/g5.bin", D.gamma5(fe));
    Fopr_DdagD op(&D);
    write_field(out + "/ddagd.bin", op.mult(fe));
    Dirac_Wilson Df(mass, &u);
    write_field(out + "/dfull.bin", Df.mult(ff));
    write_field(out + "/dfull_dag.bin", Df.mult_dag(ff));
    std::printf("IIREF_OK ops kappa=%.17g\n", D.getKappa());
    return 0;
  }

  if (cmd == "mdforce") {
    Field ee = read_field(argv[8], hsz);
    Field ze = read_field(argv[9], hsz);
    Field ef = read_field(argv[10], fsz);
    Field zf = read_field(argv[11], fsz);
    std::string out = argv[12];
    GaugeField f_eo;
    f_eo.data = D.md_force(ee, ze);
    write_field(out + "/mdf_eo.bin", f_eo.data);
    write_field(out + "/ta_eo.bin", FieldUtils::TracelessAntihermite(f_eo).data);
    Dirac_Wilson Df(mass, &u);
    write_field(out + "/mdf_full.bin", Df.md_force(ef, zf));
    Field fp(gsz), fm(gsz);
    Df.md_force_p(fp, ef, zf);
    Df.md_force_m(fm, ef, zf);
    write_field(out + "/mdf_p.bin", fp);
    write_field(out + "/mdf_m.bin", fm);
    std::printf("IIREF_OK mdforce\n");
    return 0;
  }

  double prec = std::atof(argv[9]);
  int maxiter = std::atoi(argv[10]);
  std::string out = argv[11];
  Fopr_DdagD op(&D);

  if (cmd == "cg") {
    Field b = read_field(argv[8], hsz);
    Solver_CG cg(prec, maxiter, &op);
    Field x(hsz);
    SolverOutput so = cg.solve(x, b);
    write_field(out + "/cg_x.bin", x);
    std::printf("IIREF_CG {\"iterations\": %d, \"diff\": %.17g}\n",
                so.Iterations, so.diff);
    return 0;
  }
  if (cmd == "mshift") {
    Field b = read_field(argv[8], hsz);
    std::vector<double> shifts;
    for (int i = 12; i < argc; ++i) shifts.push_back(std::atof(argv[i]));
    MultiShiftSolver_CG ms(&op, prec, maxiter);
    vector_Field xs(shifts.size());
    for (size_t i = 0; i < xs.size(); ++i) xs[i].resize(hsz);
    double diff = 0.0;
    int nconv = -1;
    SolverOutput so = ms.solve(xs, b, shifts, diff, nconv);
    for (size_t i = 0; i < xs.size(); ++i) {
      char name[64];
      std::snprintf(name, sizeof(name), "/ms_x%zu.bin", i);
      write_field(out + name, xs[i]);
    }
    std::printf("IIREF_MS {\"nconv\": %d, \"diff\": %.17g, \"out_diff\": %.17g}\n",
                nconv, diff, so.diff);
    return 0;
  }
  if (cmd == "invert") {
    Field b = read_field(argv[8], fsz);
    Solver_CG cg(prec, maxiter, &op);
    EvenOddUtils::Inverter_WilsonLike inv(&D, &cg);
    Field sol(fsz);
    inv.invert(sol, b);
    write_field(out + "/inv_x.bin", sol);
    std::printf("IIREF_INV ok\n");
    return 0;
  }
  die("unknown command");
  return 2;
}
