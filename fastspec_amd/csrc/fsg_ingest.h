// fsg_ingest.h -- host side of the line-list sidecar reader (no device code).
//
// A sidecar is a directory holding the /Lines group of a FastSpec HDF5 line list as raw
// little-endian column files (<COLUMN>.bin, one value per line, in file order) and a
// header.txt of "key=value" lines:
//   format=fastspec-lines-1
//   name=atom | <molecule name>     (the NAME attribute; 'atom...' marks an atomic list)
//   wl_vac=0|1  wmin=<A>  wmax=<A>  nl=<lines>  hflag=<int>
//   columns=W:f8,EL:S3,ION:i2,E:f4,LGGF:f4,GRAD:f4,GQST:f4,GVDW:f4[,JLO:f4,JUP:f4,GLO:f4,GUP:f4]
// Selection and conversions follow ReadHDF (src/ioutils.f90:545-1100).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace fsg {

struct SidecarCol {
  char kind = 0;  // 'f' float, 'i' integer, 'S' fixed-length string
  int size = 0;   // bytes per value
};

struct H5List;  // an open HDF5 list (fsg_hdf5.hip)

// a line list on disk: a sidecar directory, or an HDF5 file (h5 set)
struct Sidecar {
  std::string dir, name;
  std::shared_ptr<H5List> h5;
  int wl_vac = 0, hflag = 0;
  bool molecular = false;
  bool zeeman = false;  // JLO, JUP, GLO, GUP columns: a magnetic list
  bool isot = false;    // ISOT column: isotope mass numbers (0 = not given)
  double wmin = 0.0, wmax = 0.0;
  int64_t nl = 0;
  std::map<std::string, SidecarCol> cols;
};

// (element symbol, ion) -> (species slot, molecule flag) of the host's species table
using SpeciesKey = std::pair<std::string, int>;
using SpeciesMap = std::map<SpeciesKey, std::pair<int32_t, uint8_t>>;
// (element symbol, mass number) -> isotopic fraction of the host's isotope table
using IsotopeMap = std::map<SpeciesKey, double>;

// converted columns of the selected lines (the caller owns the arrays, count entries each)
struct IngestOut {
  double *wl = nullptr;
  float *elo = nullptr, *gf = nullptr, *grad = nullptr, *gqst = nullptr, *gvdw = nullptr;
  int32_t *spc = nullptr;
  uint8_t *is_mol = nullptr;
};

// Zeeman patterns of the kept lines in CSR form (sigma: blue half, then red half)
struct ZeemanOut {
  std::vector<int32_t> pi_off, si_off;
  std::vector<float> spl_pi, str_pi, spl_si, str_si;
};

// opens a sidecar directory or (by its signature) an HDF5 list file
int sidecar_open(const char *dir, Sidecar &sc, std::string &err);
// HDF5 lists through libhdf5 loaded at run time (fsg_hdf5.hip)
bool hdf5_file(const char *path);
int hdf5_open(const char *path, Sidecar &sc, std::string &err);
// rows [first, first + count) of a column, as stored (little-endian values, fixed strings)
int hdf5_read(const Sidecar &sc, const std::string &name, int64_t first, int64_t count, void *dst, std::string &err);
// lines [first, first + count) of the file are inside [wstart, wend] (0-based first)
int sidecar_select(const Sidecar &sc, double wstart, double wend, int64_t &first, int64_t &count, std::string &err);
// reads and converts lines [first, first + count) with nthreads threads into out (packed,
// lines above the EOS ion limit dropped); kept receives the number written
int sidecar_convert(const Sidecar &sc, int64_t first, int64_t count, const SpeciesMap &species,
                    const IsotopeMap &isotopes, int32_t mol_slot, int nthreads, const IngestOut &out, int64_t &kept,
                    std::string &err, ZeemanOut *z = nullptr);
// ZeemanPattern (src/utils.f90:901-972 with MergeShifts 976-1004): pi shifts and strengths,
// one sigma branch
void zeeman_pattern(float jlo, float jup, float glo, float gup, std::vector<float> &spi, std::vector<float> &rpi,
                    std::vector<float> &ssig, std::vector<float> &rsig);
double vac2air(double w);

}  // namespace fsg
