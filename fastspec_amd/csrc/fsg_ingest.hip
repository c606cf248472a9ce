// fsg_ingest.hip -- the line-list sidecar reader (host code only).
//
// Reference routines restated here (file:line in the reference tree):
//   ReadHDF: range cull, W read, sort check, bounds, column slices, ion filter, isotopic
//            fractions and the conversions of gf and the damping constants
//                                                          src/ioutils.f90:696-1076
//   LowerBoundReal8, UpperBoundReal8                       src/utils.f90:701-737
//   Vac2Air                                                src/utils.f90:1007-1017
//   ZeemanPattern, MergeShifts; the reader's pi/sigma layout src/utils.f90:901-1004,
//                                                          src/ioutils.f90:997-1031
// The columns are read with pread from raw files by several threads (the reference reads
// HDF5 hyperslabs); every other step keeps the reference's order and REAL(4) arithmetic.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <thread>

#include "fsg_ingest.h"

namespace fsg {

namespace {

constexpr int kMaxIon = 5;  // Maxion, src/shared.f90:434

// reads count values of size bytes from column file <dir>/<name>.bin starting at value first
// (or rows of the HDF5 list's column dataset)
bool read_col(const Sidecar &sc, const std::string &name, int64_t first, int64_t count, void *dst, std::string &err) {
  if (sc.h5) return hdf5_read(sc, name, first, count, dst, err) == 0;
  auto it = sc.cols.find(name);
  if (it == sc.cols.end()) {
    err = "fsg_lines: column " + name + " missing in " + sc.dir;
    return false;
  }
  const int sz = it->second.size;
  const std::string fn = sc.dir + "/" + name + ".bin";
  const int fd = ::open(fn.c_str(), O_RDONLY);
  if (fd < 0) {
    err = "fsg_lines: cannot open " + fn;
    return false;
  }
  char *p = static_cast<char *>(dst);
  int64_t off = first * sz, left = count * sz;
  while (left > 0) {
    const ssize_t r = ::pread(fd, p, (size_t)std::min<int64_t>(left, 1LL << 30), (off_t)off);
    if (r <= 0) {
      ::close(fd);
      err = "fsg_lines: short read of " + fn;
      return false;
    }
    p += r;
    off += r;
    left -= r;
  }
  ::close(fd);
  return true;
}

std::string trim_el(const char *s, int n) {  // ADJUSTL(TRIM(El)): blanks and NULs removed
  int a = 0, b = n;
  while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\0')) b--;
  while (a < b && s[a] == ' ') a++;
  return std::string(s + a, s + b);
}

// REAL(4) 10.0**x as the reference's compiler evaluates it: the power of ten becomes the C
// library's exp10f (the same bits as flang's scalar code; powf differs in about one value
// in five by one ulp)
inline float pow10f(float x) { return ::exp10f(x); }

// components coinciding in shift within tol become one with the summed strength
void merge_shifts(std::vector<double> &g, std::vector<double> &r, double tol) {
  std::vector<double> gm, rm;
  for (size_t i = 0; i < g.size(); i++) {
    size_t k = 0;
    for (; k < gm.size(); k++)
      if (std::fabs(g[i] - gm[k]) < tol) {
        rm[k] = rm[k] + r[i];
        break;
      }
    if (k == gm.size()) {
      gm.push_back(g[i]);
      rm.push_back(r[i]);
    }
  }
  g.swap(gm);
  r.swap(rm);
}

}  // namespace

void zeeman_pattern(float jlo, float jup, float glo, float gup, std::vector<float> &spi, std::vector<float> &rpi,
                    std::vector<float> &ssig, std::vector<float> &rsig) {
  const double tol = 1e-5;
  const double Jl = jlo, Ju = jup, Gl = glo, Gu = gup;
  const int nm = (int)std::lround(2.0 * Ju) + 1;
  std::vector<double> gp, rp, gs, rs;
  for (int i = 1; i <= nm; i++) {  // upper sublevels m = Jup .. -Jup
    const double m = Ju - (double)(i - 1);
    double p, q;
    if (Jl == Ju) {
      p = m * m;
      q = 0.25 * (Ju - m) * (Ju + m + 1.0);
    } else if (Jl > Ju) {
      p = (Ju + 1.0) * (Ju + 1.0) - m * m;
      q = 0.25 * (Ju + m + 1.0) * (Ju + m + 2.0);
    } else {
      p = Ju * Ju - m * m;
      q = 0.25 * (Ju - m) * (Ju - m - 1.0);
    }
    double a = m * (Gu - Gl), b = m * Gu - (m + 1.0) * Gl;
    if (std::fabs(a) < tol) a = 0.0;
    if (std::fabs(b) < tol) b = 0.0;
    gp.push_back(a);
    rp.push_back(p);
    gs.push_back(b);
    rs.push_back(q);
  }
  std::vector<double> g1, r1, g2, r2;  // vanishing intensities dropped; sigma sign convention
  for (int i = 0; i < nm; i++) {
    if (std::fabs(rp[i]) > tol) {
      g1.push_back(gp[i]);
      r1.push_back(rp[i]);
    }
    if (std::fabs(rs[i]) > tol) {
      g2.push_back(-gs[i]);
      r2.push_back(rs[i]);
    }
  }
  merge_shifts(g1, r1, tol);
  merge_shifts(g2, r2, tol);
  double s1 = 0.0, s2 = 0.0;
  for (double v : r1) s1 = s1 + v;
  for (double v : r2) s2 = s2 + v;
  spi.clear();
  rpi.clear();
  ssig.clear();
  rsig.clear();
  for (size_t i = 0; i < g1.size(); i++) {
    spi.push_back((float)g1[i]);
    rpi.push_back((float)(s1 > 0.0 ? r1[i] / s1 : r1[i]));
  }
  for (size_t i = 0; i < g2.size(); i++) {
    ssig.push_back((float)g2[i]);
    rsig.push_back((float)(s2 > 0.0 ? r2[i] / s2 : r2[i]));
  }
}

double vac2air(double w) {
  const double S = 1e4 / w, S2 = S * S;
  const double N = 1.0 + 8.34254e-5 + 2.406147e-2 / (130.0 - S2) + 1.5998e-4 / (38.9 - S2);
  return w > 2000.0 ? w / N : w;
}

namespace {

// the column set and types a list must have (both sources)
int check_columns(Sidecar &sc, std::string &err);

}  // namespace

int sidecar_open(const char *dir, Sidecar &sc, std::string &err) {
  sc = Sidecar();
  sc.dir = dir ? dir : "";
  if (dir && hdf5_file(dir)) {
    if (hdf5_open(dir, sc, err)) return 1;
    return check_columns(sc, err);
  }
  std::ifstream in(sc.dir + "/header.txt");
  if (!in) {
    err = "fsg_lines: " + sc.dir + " has no header.txt";
    return 1;
  }
  std::map<std::string, std::string> kv;
  std::string line;
  while (std::getline(in, line)) {
    const size_t e = line.find('=');
    if (e == std::string::npos) continue;
    kv[line.substr(0, e)] = line.substr(e + 1);
  }
  if (kv["format"] != "fastspec-lines-1") {
    err = "fsg_lines: " + sc.dir + " is not a FastSpec line-list sidecar";
    return 1;
  }
  try {
    sc.name = kv["name"];
    sc.wl_vac = std::stoi(kv.count("wl_vac") ? kv["wl_vac"] : "0");
    sc.hflag = std::stoi(kv.count("hflag") ? kv["hflag"] : "0");
    sc.wmin = std::stod(kv.at("wmin"));
    sc.wmax = std::stod(kv.at("wmax"));
    sc.nl = std::stoll(kv.at("nl"));
    std::stringstream cs(kv.at("columns"));
    std::string item;
    while (std::getline(cs, item, ',')) {
      const size_t c = item.find(':');
      if (c == std::string::npos || c + 2 > item.size()) throw std::invalid_argument("column");
      SidecarCol col;
      col.kind = item[c + 1];
      col.size = std::stoi(item.substr(c + 2));
      if ((col.kind != 'f' && col.kind != 'i' && col.kind != 'S') || col.size <= 0) throw std::invalid_argument("type");
      sc.cols[item.substr(0, c)] = col;
    }
  } catch (...) {
    err = "fsg_lines: malformed header.txt in " + sc.dir;
    return 1;
  }
  sc.molecular = sc.name.compare(0, 4, "atom") != 0;
  return check_columns(sc, err);
}

namespace {

int check_columns(Sidecar &sc, std::string &err) {
  const std::vector<std::string> need =
      sc.molecular ? std::vector<std::string>{"W", "E", "LGGF", "GRAD"}
                   : std::vector<std::string>{"W", "EL", "ION", "E", "LGGF", "GRAD", "GQST", "GVDW"};
  for (const auto &n : need)
    if (!sc.cols.count(n)) {
      err = "fsg_lines: column " + n + " missing in " + sc.dir;
      return 1;
    }
  const auto ck = [&](const char *n, char k, int s) {
    const SidecarCol &c = sc.cols.at(n);
    return c.kind == k && c.size == s;
  };
  bool ok = ck("W", 'f', 8) && ck("E", 'f', 4) && ck("LGGF", 'f', 4) && ck("GRAD", 'f', 4);
  if (!sc.molecular)
    ok = ok && ck("GQST", 'f', 4) && ck("GVDW", 'f', 4) && sc.cols.at("ION").kind == 'i' &&
         (sc.cols.at("ION").size == 2 || sc.cols.at("ION").size == 4) && sc.cols.at("EL").kind == 'S';
  if (!ok) {
    err = "fsg_lines: unsupported column types in " + sc.dir;
    return 1;
  }
  sc.isot = sc.cols.count("ISOT") != 0;
  if (sc.isot && (sc.molecular || sc.cols.at("ISOT").kind != 'i' ||
                  (sc.cols.at("ISOT").size != 2 && sc.cols.at("ISOT").size != 4))) {
    err = "fsg_lines: unsupported ISOT column in " + sc.dir;
    return 1;
  }
  sc.zeeman = sc.cols.count("JLO") != 0;
  if (sc.zeeman &&
      (sc.molecular || !sc.cols.count("JUP") || !sc.cols.count("GLO") || !sc.cols.count("GUP") ||
       !ck("JLO", 'f', 4) || !ck("JUP", 'f', 4) || !ck("GLO", 'f', 4) || !ck("GUP", 'f', 4))) {
    err = "fsg_lines: incomplete Zeeman columns in " + sc.dir;
    return 1;
  }
  return 0;
}

}  // namespace

int sidecar_select(const Sidecar &sc, double wstart, double wend, int64_t &first, int64_t &count, std::string &err) {
  first = 0;
  count = 0;
  // whole-file cull on the WMIN/WMAX attributes, converted to air first (ioutils.f90:696-727)
  double wmin = sc.wmin, wmax = sc.wmax;
  if (sc.wl_vac == 1) {
    wmin = vac2air(wmin);
    wmax = vac2air(wmax);
  }
  if (wmax < wstart || wmin > wend || sc.nl <= 0) return 0;
  // the full W column, converted, checked for order (ioutils.f90:729-749)
  std::vector<double> W((size_t)sc.nl);
  if (!read_col(sc, "W", 0, sc.nl, W.data(), err)) return 1;
  if (sc.wl_vac == 1)
    for (double &w : W) w = vac2air(w);
  for (int64_t i = 1; i < sc.nl; i++)
    if (!(W[i - 1] <= W[i])) {
      err = "ReadHDF: wavelength array is not sorted in " + sc.dir;
      return 1;
    }
  // LowerBoundReal8 / UpperBoundReal8 (utils.f90:701-737), 1-based I1, I2
  const int64_t i1 = (int64_t)(std::lower_bound(W.begin(), W.end(), wstart) - W.begin()) + 1;
  const int64_t i2 = (int64_t)(std::upper_bound(W.begin(), W.end(), wend) - W.begin());
  if (i1 > i2 || i1 > sc.nl || i2 < 1) return 0;
  first = i1 - 1;
  count = i2 - i1 + 1;
  return 0;
}

int sidecar_convert(const Sidecar &sc, int64_t first, int64_t count, const SpeciesMap &species,
                    const IsotopeMap &isotopes, int32_t mol_slot, int nthreads, const IngestOut &out, int64_t &kept,
                    std::string &err, ZeemanOut *z) {
  kept = 0;
  if (z) {
    *z = ZeemanOut();
    z->pi_off.push_back(0);
    z->si_off.push_back(0);
  }
  if (count <= 0) return 0;
  if (sc.zeeman && !z) {
    err = "fsg_lines: " + sc.dir + " is a Zeeman list: read it with fsg_lines_stream";
    return 1;
  }
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(std::max(nthreads, 1), (count + 65535) / 65536));
  struct Part {
    int64_t a = 0, n = 0, kept = 0;
    std::vector<double> wl;
    std::vector<float> elo, gf, grad, gqst, gvdw;
    std::vector<int32_t> spc;
    std::vector<uint8_t> mol;
    ZeemanOut zp;  // this part's patterns, offsets from 0
    std::string err;
  };
  std::vector<Part> parts((size_t)nt);
  for (int t = 0; t < nt; t++) {
    parts[t].a = first + count * t / nt;
    parts[t].n = first + count * (t + 1) / nt - parts[t].a;
  }
  const int elsz = sc.molecular ? 0 : sc.cols.at("EL").size;
  const int ionsz = sc.molecular ? 0 : sc.cols.at("ION").size;
  auto work = [&](Part &P) {
    const int64_t n = P.n;
    std::vector<double> W((size_t)n);
    std::vector<float> E((size_t)n), G((size_t)n), R((size_t)n), Q, V;
    std::vector<char> EL;
    std::vector<char> ION;
    bool ok = read_col(sc, "W", P.a, n, W.data(), P.err) && read_col(sc, "E", P.a, n, E.data(), P.err) &&
              read_col(sc, "LGGF", P.a, n, G.data(), P.err) && read_col(sc, "GRAD", P.a, n, R.data(), P.err);
    if (ok && !sc.molecular) {
      Q.resize((size_t)n);
      V.resize((size_t)n);
      EL.resize((size_t)n * elsz);
      ION.resize((size_t)n * ionsz);
      ok = read_col(sc, "GQST", P.a, n, Q.data(), P.err) && read_col(sc, "GVDW", P.a, n, V.data(), P.err) &&
           read_col(sc, "EL", P.a, n, EL.data(), P.err) && read_col(sc, "ION", P.a, n, ION.data(), P.err);
    }
    std::vector<char> ISO;
    const int isosz = sc.isot ? sc.cols.at("ISOT").size : 0;
    if (ok && sc.isot) {
      ISO.resize((size_t)n * isosz);
      ok = read_col(sc, "ISOT", P.a, n, ISO.data(), P.err);
    }
    std::vector<float> JL, JU, GL, GU;
    if (ok && sc.zeeman) {
      JL.resize((size_t)n);
      JU.resize((size_t)n);
      GL.resize((size_t)n);
      GU.resize((size_t)n);
      ok = read_col(sc, "JLO", P.a, n, JL.data(), P.err) && read_col(sc, "JUP", P.a, n, JU.data(), P.err) &&
           read_col(sc, "GLO", P.a, n, GL.data(), P.err) && read_col(sc, "GUP", P.a, n, GU.data(), P.err);
      P.zp.pi_off.push_back(0);
      P.zp.si_off.push_back(0);
    }
    if (!ok) return;
    std::vector<float> spi, rpi, ssig, rsig;
    P.wl.reserve((size_t)n);
    P.elo.reserve((size_t)n);
    P.gf.reserve((size_t)n);
    P.grad.reserve((size_t)n);
    P.spc.reserve((size_t)n);
    if (!sc.molecular) {
      P.gqst.reserve((size_t)n);
      P.gvdw.reserve((size_t)n);
      P.mol.reserve((size_t)n);
    }
    for (int64_t i = 0; i < n; i++) {
      double w = W[(size_t)i];
      if (sc.wl_vac == 1) w = vac2air(w);
      if (sc.molecular) {  // LinesM: Grad as stored, gf = 10**LGGF (ioutils.f90:806-813)
        P.wl.push_back(w);
        P.elo.push_back(E[(size_t)i]);
        P.gf.push_back(pow10f(G[(size_t)i]));
        P.grad.push_back(R[(size_t)i]);
        P.spc.push_back(mol_slot);
        continue;
      }
      const int ion = ionsz == 2 ? (int)reinterpret_cast<const int16_t *>(ION.data())[i]
                                 : (int)reinterpret_cast<const int32_t *>(ION.data())[i];
      if (ion > kMaxIon) continue;  // ions above the EOS limit (ioutils.f90:921-923)
      const std::string el = trim_el(EL.data() + (size_t)i * elsz, elsz);
      auto it = species.find(SpeciesKey(el, ion));
      if (it == species.end()) {
        P.err = "fsg_lines: species " + el + " " + std::to_string(ion) + " of " + sc.dir +
                " is not in the species table";
        return;
      }
      // isotopic fraction of a line with a mass number (ioutils.f90:940-962), 1 otherwise
      float giso = 1.0f;
      if (sc.isot) {
        const int a = isosz == 2 ? (int)reinterpret_cast<const int16_t *>(ISO.data())[i]
                                 : (int)reinterpret_cast<const int32_t *>(ISO.data())[i];
        if (a > 0) {
          auto f = isotopes.find(SpeciesKey(el, a));
          if (f == isotopes.end()) {
            P.err = "ReadHDF: isotope " + std::to_string(a) + el + " in the ISOT column of " + sc.dir +
                    " is not in the isotopic table";
            return;
          }
          giso = (float)f->second;
        }
      }
      // LinesA (ioutils.f90:1056-1076): gf = 10**LGGF (times the isotopic fraction); GRAD
      // 10** when 0 < x < 20; GQST and GVDW 10** when negative
      const float g = R[(size_t)i], q = Q[(size_t)i], v = V[(size_t)i];
      P.wl.push_back(w);
      P.elo.push_back(E[(size_t)i]);
      P.gf.push_back(sc.isot ? pow10f(G[(size_t)i]) * giso : pow10f(G[(size_t)i]));
      P.grad.push_back((g < 20.0f && g > 0.0f) ? pow10f(g) : g);
      P.gqst.push_back(q < 0.0f ? pow10f(q) : q);
      P.gvdw.push_back(v < 0.0f ? pow10f(v) : v);
      P.spc.push_back(it->second.first);
      P.mol.push_back(it->second.second);
      if (sc.zeeman) {
        ZeemanOut &Z = P.zp;
        if (GL[(size_t)i] > 90.0f || GU[(size_t)i] > 90.0f) {
          // a missing Lande factor: one unshifted pi component and a null sigma pair
          Z.spl_pi.push_back(0.0f);
          Z.str_pi.push_back(1.0f);
          Z.spl_si.insert(Z.spl_si.end(), {0.0f, 0.0f});
          Z.str_si.insert(Z.str_si.end(), {1.0f, 1.0f});
        } else {
          zeeman_pattern(JL[(size_t)i], JU[(size_t)i], GL[(size_t)i], GU[(size_t)i], spi, rpi, ssig, rsig);
          Z.spl_pi.insert(Z.spl_pi.end(), spi.begin(), spi.end());
          Z.str_pi.insert(Z.str_pi.end(), rpi.begin(), rpi.end());
          for (size_t k = 0; k < ssig.size(); k++) {  // blue (negative shift) half
            Z.spl_si.push_back(-ssig[k]);
            Z.str_si.push_back(rsig[k]);
          }
          for (size_t k = 0; k < ssig.size(); k++) {  // red half
            Z.spl_si.push_back(ssig[k]);
            Z.str_si.push_back(rsig[k]);
          }
        }
        Z.pi_off.push_back((int32_t)Z.spl_pi.size());
        Z.si_off.push_back((int32_t)Z.spl_si.size());
      }
    }
    P.kept = (int64_t)P.wl.size();
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; t++) th.emplace_back(work, std::ref(parts[t]));
  work(parts[0]);
  for (auto &x : th) x.join();
  for (const Part &P : parts)
    if (!P.err.empty()) {
      err = P.err;
      return 1;
    }
  int64_t o = 0;
  for (const Part &P : parts) {
    const size_t k = (size_t)P.kept;
    std::memcpy(out.wl + o, P.wl.data(), k * sizeof(double));
    std::memcpy(out.elo + o, P.elo.data(), k * sizeof(float));
    std::memcpy(out.gf + o, P.gf.data(), k * sizeof(float));
    std::memcpy(out.grad + o, P.grad.data(), k * sizeof(float));
    std::memcpy(out.spc + o, P.spc.data(), k * sizeof(int32_t));
    if (!sc.molecular) {
      if (out.gqst) std::memcpy(out.gqst + o, P.gqst.data(), k * sizeof(float));
      if (out.gvdw) std::memcpy(out.gvdw + o, P.gvdw.data(), k * sizeof(float));
      if (out.is_mol) std::memcpy(out.is_mol + o, P.mol.data(), k);
    }
    o += P.kept;
    if (z && sc.zeeman) {  // concatenate the parts' patterns, offsets shifted
      const int32_t bp = (int32_t)z->spl_pi.size(), bs = (int32_t)z->spl_si.size();
      for (size_t k = 1; k < P.zp.pi_off.size(); k++) {
        z->pi_off.push_back(bp + P.zp.pi_off[k]);
        z->si_off.push_back(bs + P.zp.si_off[k]);
      }
      z->spl_pi.insert(z->spl_pi.end(), P.zp.spl_pi.begin(), P.zp.spl_pi.end());
      z->str_pi.insert(z->str_pi.end(), P.zp.str_pi.begin(), P.zp.str_pi.end());
      z->spl_si.insert(z->spl_si.end(), P.zp.spl_si.begin(), P.zp.spl_si.end());
      z->str_si.insert(z->str_si.end(), P.zp.str_si.begin(), P.zp.str_si.end());
    }
  }
  kept = o;
  return 0;
}

}  // namespace fsg
