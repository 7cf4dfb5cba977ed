// The reference's double-precision dump and restart files, host side.
//
//   data.NNNN.dbl  each dumped variable's active zones in turn, i fastest,
//                  native-endian doubles (Src/write_data.cpp:136-159,
//                  FillOutputBuffer :564-607)
//   dbl.out        one text line per dump: nfile, time, dt, step counter,
//                  "single_file", the endianness, the variable names
//                  (Src/write_data.cpp:476-505)
//   restart.out    one binary Restart record per dump
//                  (Src/structs.hpp:382-390, RestartDump Src/restart.cpp:417-453)
#include "dbl_io.hpp"

#include "../../include/gpluto_amd.h"

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <utility>

namespace gpl {
namespace {

constexpr int kDblOutput = 1;        // DBL_OUTPUT: its slot in the output table (Src/gpluto.hpp:139,
                                     // Src/runtime_setup.cpp:252)
constexpr int kOutputTypes = 16;     // MAX_OUTPUT_TYPES (Src/gpluto.hpp:184)
constexpr int kRestartParams = 12;   // MAX_RESTART_PARAMS (Src/gpluto.hpp:186)

// restart.out record: the reference's Restart, member for member, so the
// files interchange with its builds on the same ABI (x86-64: 256 bytes)
struct RestartRecord {
  double t;
  double dt;
  double param[kRestartParams];
  int nstep;
  int nfile[kOutputTypes];
  int nfile_particles[kOutputTypes];
  char fill[10];
};
static_assert(sizeof(RestartRecord) == 256, "restart.out records are 256 bytes");
static_assert(offsetof(RestartRecord, nstep) == 112 && offsetof(RestartRecord, nfile) == 116,
              "restart.out member offsets");

bool host_little_endian()
{
  const uint16_t one = 1;
  unsigned char b;
  memcpy(&b, &one, 1);
  return b == 1;
}

template <typename T>
void swap_bytes(T &v)
{
  unsigned char b[sizeof(T)];
  memcpy(b, &v, sizeof(T));
  for (size_t i = 0; i < sizeof(T)/2; i++) std::swap(b[i], b[sizeof(T) - 1 - i]);
  memcpy(&v, b, sizeof(T));
}

std::string join(const char *dir, const std::string &name) { return std::string(dir) + "/" + name; }

std::string dump_name(int nfile)
{
  char b[32];
  snprintf(b, sizeof(b), "data.%04d.dbl", nfile);
  return b;
}

// Line `index` of a text file, without its newline; false when absent.
bool read_line(const std::string &path, int index, std::string &line, bool &opened)
{
  FILE *f = fopen(path.c_str(), "r");
  opened = f != nullptr;
  if (!f) return false;
  std::string cur;
  int l = 0, ch;
  bool found = false;
  while ((ch = fgetc(f)) != EOF) {
    if (ch == '\n') {
      if (l == index) { found = true; break; }
      l++;
      cur.clear();
    } else {
      cur.push_back((char)ch);
    }
  }
  if (!found && l == index && !cur.empty()) found = true;   // last line without a newline
  fclose(f);
  if (found) line = cur;
  return found;
}

std::vector<std::string> tokens(const std::string &s)
{
  std::vector<std::string> t;
  size_t i = 0;
  while (i < s.size()) {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\r')) i++;
    size_t j = i;
    while (j < s.size() && s[j] != ' ' && s[j] != '\t' && s[j] != '\r') j++;
    if (j > i) t.push_back(s.substr(i, j - i));
    i = j;
  }
  return t;
}

}  // namespace

int dbl_write(const char *dir, int nfile, const DumpStep &st, const std::vector<std::string> &names,
              const double *data, size_t count)
{
  if (!dir || nfile < 0 || (!data && count)) return GPL_EINVAL;

  // 1. the dump: a fresh file per nfile, the old one deleted first
  //    (FileDelete, Src/write_data.cpp:141), which also spares the kernel
  //    writing back the old file's pages
  const std::string path = join(dir, dump_name(nfile));
  std::remove(path.c_str());
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return GPL_EIO;
  const size_t w = count ? fwrite(data, sizeof(double), count, f) : 0;
  if (fclose(f) != 0 || w != count) return GPL_EIO;

  // 2. dbl.out: created with dump 0, else line nfile is written after the
  //    first nfile lines, in place (Src/write_data.cpp:476-484)
  f = fopen(join(dir, "dbl.out").c_str(), nfile == 0 ? "w" : "r+");
  if (!f) return GPL_EIO;
  for (int l = 0; l < nfile; l++) {
    int ch;
    while ((ch = fgetc(f)) != EOF && ch != '\n') {}
    if (ch == EOF) break;
  }
  fseek(f, ftell(f), SEEK_SET);   // a read followed by a write needs a seek
  fprintf(f, "%d %12.6e %12.6e %d ", nfile, st.t, st.dt, st.nstep);
  fprintf(f, "single_file ");
  fprintf(f, host_little_endian() ? "little " : "big ");
  for (const std::string &n : names) fprintf(f, "%s ", n.c_str());
  fprintf(f, "\n");
  if (fclose(f) != 0) return GPL_EIO;

  // 3. restart.out: record nfile (the reference writes its dumps' records in
  //    order, so record k is dump k); members the reference leaves unset
  //    (nfileParticles, fill) as zeros
  RestartRecord r;
  memset(&r, 0, sizeof(r));
  r.t = st.t;
  r.dt = st.dt;
  r.nstep = st.nstep;
  for (int q = 0; q < kOutputTypes; q++) r.nfile[q] = -1;   // never written (Src/set_output.cpp:106)
  r.nfile[kDblOutput] = nfile;
  f = fopen(join(dir, "restart.out").c_str(), nfile == 0 ? "wb" : "r+b");
  if (!f) return GPL_EIO;
  if (fseek(f, (long)nfile*(long)sizeof(r), SEEK_SET) != 0 || fwrite(&r, sizeof(r), 1, f) != 1) {
    fclose(f);
    return GPL_EIO;
  }
  return fclose(f) != 0 ? GPL_EIO : 0;
}

int dbl_read(const char *dir, int nfile, const std::vector<std::string> &names, size_t count,
             DumpStep *st, std::vector<double> &data)
{
  // a negative file number (count from the end) is disabled in the
  // reference too (Src/restart.cpp:350-352)
  if (!dir || !st || nfile < 0) return GPL_EINVAL;

  // 1. dbl.out line nfile: single_file mode, endianness, the variables
  //    (Src/restart.cpp:69-106)
  std::string line;
  bool opened = false;
  if (!read_line(join(dir, "dbl.out"), nfile, line, opened)) return opened ? GPL_EINVAL : GPL_EIO;
  const std::vector<std::string> t = tokens(line);
  if (t.size() < 6) return GPL_EINVAL;
  if (t[4] != "single_file") return GPL_EUNSUPPORTED;
  if (t[5] != "little" && t[5] != "big") return GPL_EINVAL;
  const bool swap = (t[5] == "big") == host_little_endian();
  if (t.size() - 6 != names.size()) return GPL_EINVAL;
  for (size_t q = 0; q < names.size(); q++)
    if (t[6 + q] != names[q]) return GPL_EINVAL;

  // 2. restart.out: the first record of dump nfile (Src/restart.cpp:356-392)
  FILE *f = fopen(join(dir, "restart.out").c_str(), "rb");
  if (!f) return GPL_EIO;
  RestartRecord r;
  bool hit = false;
  while (fread(&r, sizeof(r), 1, f) == 1) {
    int nf = r.nfile[kDblOutput];
    if (swap) swap_bytes(nf);
    if (nf == nfile) { hit = true; break; }
  }
  fclose(f);
  if (!hit) return GPL_EINVAL;
  if (swap) { swap_bytes(r.t); swap_bytes(r.dt); swap_bytes(r.nstep); }
  st->t = r.t;
  st->dt = r.dt;
  st->nstep = r.nstep;

  // 3. the dump: exactly `count` doubles
  f = fopen(join(dir, dump_name(nfile)).c_str(), "rb");
  if (!f) return GPL_EIO;
  if (fseek(f, 0, SEEK_END) != 0) { fclose(f); return GPL_EIO; }
  const long bytes = ftell(f);
  if (bytes < 0 || (size_t)bytes != count*sizeof(double)) { fclose(f); return GPL_EINVAL; }
  rewind(f);
  data.resize(count);
  const size_t got = count ? fread(data.data(), sizeof(double), count, f) : 0;
  fclose(f);
  if (got != count) return GPL_EIO;
  if (swap)
    for (double &v : data) swap_bytes(v);
  return 0;
}

}  // namespace gpl
