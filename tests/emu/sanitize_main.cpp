// sanitize_main.cpp -- TEST INFRASTRUCTURE ONLY.
//
// The emulated column kernels (emu_vert.cpp) as a program of their own, so
// that the host sanitizers can watch them: ThreadSanitizer reports an LDS
// location one lane writes and another lane reads (or writes) with no
// barrier ordering the two; AddressSanitizer reports an index outside the
// LDS block or scratch a launch gets.  Built by tests/test_emu_sanitizers.py
// with g++ only.
//
//   sanitize_main <case file>
//
// Case file (native little-endian): int32 hdr[8] = {magic 1196840532,
// n_cols, n_alts, solver, n_steps, flags, waves, per_wg}, flags bit 0 ions,
// 1 UseMSIS, 2 GSWM tides, 3 WACCM tides, 4 UseDamping, 5 carries in LDS for
// the slot run, 6 the GPU's forms of the arithmetic (device_math); then float64 [col][component][level] arrays: nds(6) vvel(6)
// vel(3) temp rho tempunit mmm gamma cmax eddy alt grav invr dalt, coslat
// [col]; with ions ions(10) ivel(3) sza[col] mlat[col]; with UseMSIS
// msis[col][9]; with tides tides[col][3][2].
//
// Runs the columns as whole workgroups (gv_emu_advance, `waves` wavefronts)
// and, when per_wg > 0, again through the short-column slots of the
// ensemble step (gv_emu_advance_slots, per_wg columns per block); the two
// must give the same bits.  Exit status 0 and "ok" on success.
#include <stdio.h>
#include <string.h>

#define GV_EMU_FORWARD_ONLY 1
#include "emu_vert.cpp"

namespace {

gitm_vert_field level_field(double *p, int ncomp, int L) {
  gitm_vert_field f{};
  f.ptr = p;
  f.s_lon = (long long)ncomp * L;
  f.s_lev = 1;
  f.s_comp = L;
  f.first_level = -1;
  return f;
}

gitm_vert_field column_field(double *p, int ncomp) {
  gitm_vert_field f{};
  f.ptr = p;
  f.s_lon = ncomp;
  f.s_comp = 1;
  f.first_level = 0;
  return f;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: %s case-file\n", argv[0]);
    return 2;
  }
  FILE *fp = fopen(argv[1], "rb");
  if (!fp) return 2;
  int hdr[8];
  if (fread(hdr, sizeof(int), 8, fp) != 8 || hdr[0] != 1196840532) return 2;
  const int nc = hdr[1], na = hdr[2], solver = hdr[3], n_steps = hdr[4], flags = hdr[5];
  const int waves = hdr[6], per_wg = hdr[7], L = na + 4;
  const bool ions = flags & 1, msis = flags & 2, gswm = flags & 4, waccm = flags & 8;
  struct Arr {
    const char *name;
    int ncomp;    // components per column (level fields: times L)
    bool level;   // one value per level
    bool state;   // written by the step
    bool present;
  };
  Arr arrs[] = {{"nds", 6, true, true, true},        {"vvel", 6, true, true, true},
                {"vel", 3, true, true, true},        {"temp", 1, true, true, true},
                {"rho", 1, true, true, true},        {"tempunit", 1, true, false, true},
                {"mmm", 1, true, false, true},       {"gamma", 1, true, false, true},
                {"cmax", 1, true, false, true},      {"eddy", 1, true, false, true},
                {"alt", 1, true, false, true},       {"grav", 1, true, false, true},
                {"invr", 1, true, false, true},      {"dalt", 1, true, false, true},
                {"coslat", 1, false, false, true},   {"ions", 10, true, true, ions},
                {"ivel", 3, true, false, ions},      {"sza", 1, false, false, ions},
                {"mlat", 1, false, false, ions},     {"msis", 9, false, false, msis},
                {"tides", 6, false, false, gswm || waccm}};
  const int na_ = (int)(sizeof(arrs) / sizeof(arrs[0]));
  std::vector<std::vector<double>> in(na_);
  for (int i = 0; i < na_; i++) {
    if (!arrs[i].present) continue;
    const size_t n = (size_t)nc * arrs[i].ncomp * (arrs[i].level ? L : 1);
    in[i].resize(n);
    if (fread(in[i].data(), sizeof(double), n, fp) != n) return 2;
  }
  fclose(fp);

  gitm_vert_params p;
  gv::default_params(&p);
  p.use_ion_advection = ions;
  p.use_msis = msis;
  p.use_gswm_tides = gswm;
  p.use_waccm_tides = waccm;
  p.use_damping = (flags & 16) != 0;

  // one run over copies of the inputs; returns the state after it
  auto run = [&](bool slots) {
    std::vector<std::vector<double>> a = in;
    std::vector<double> logns((size_t)nc * 6 * L), nden((size_t)nc * L);
    gitm_vert_columns c{};
    c.n_lons = nc;
    c.n_lats = 1;
    c.n_blocks = 1;
    c.n_alts = na;
    gitm_vert_field *lf[] = {&c.ndensity_s, &c.vertical_velocity, &c.velocity, &c.temperature,
                             &c.rho, &c.temp_unit, &c.mean_major_mass, &c.gamma, &c.cmax,
                             &c.eddy, &c.altitude, &c.gravity, &c.inv_radial_distance, &c.dalt};
    for (int i = 0; i < 14; i++) *lf[i] = level_field(a[i].data(), arrs[i].ncomp, L);
    c.cos_latitude = column_field(a[14].data(), 1);
    c.log_ns = level_field(logns.data(), 6, L);
    c.ndensity = level_field(nden.data(), 1, L);
    if (ions) {
      c.idensity_s = level_field(a[15].data(), 10, L);
      c.ivelocity = level_field(a[16].data(), 3, L);
      c.sza = column_field(a[17].data(), 1);
      c.mlat = column_field(a[18].data(), 1);
    }
    if (msis) c.msis_bc = column_field(a[19].data(), 9);
    if (gswm || waccm) {
      gitm_vert_field *tf[] = {&c.tides_east, &c.tides_north, &c.tides_temp};
      for (int i = 0; i < 3; i++) {
        gitm_vert_field f{};
        f.ptr = a[20].data() + 2 * i;
        f.s_lon = 6;
        f.s_lev = 1;
        f.first_level = -1;
        *tf[i] = f;
      }
    }
    gitm_vert_error e{};
    const int dm = (flags & 64) != 0;
    const int rc = slots ? gv_emu_advance_slots(&p, &c, solver, n_steps, per_wg,
                                                (flags & 32) != 0, 0, nc, dm, &e)
                         : gv_emu_advance(&p, &c, solver, n_steps, waves, 0, nc, dm, &e);
    if (rc) {
      fprintf(stderr, "%s run: error %d at column %d level %d\n", slots ? "slot" : "workgroup",
              rc, e.column, e.level);
      exit(1);
    }
    a.push_back(logns);
    a.push_back(nden);
    return a;
  };

  const std::vector<std::vector<double>> ref = run(false);
  if (per_wg > 0) {
    const std::vector<std::vector<double>> got = run(true);
    for (size_t i = 0; i < ref.size(); i++)
      if (ref[i].size() != got[i].size() ||
          (!ref[i].empty() &&
           memcmp(ref[i].data(), got[i].data(), ref[i].size() * sizeof(double)) != 0)) {
        fprintf(stderr, "slot run differs from the workgroup run in array %zu\n", i);
        return 1;
      }
  }
  printf("ok\n");
  return 0;
}
