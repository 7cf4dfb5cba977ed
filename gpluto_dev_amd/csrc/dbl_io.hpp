// The reference's double-precision dump and restart files, host side:
// data.NNNN.dbl (single_file mode), dbl.out and restart.out.
#pragma once
#include <cstddef>
#include <string>
#include <vector>

namespace gpl {

// Step state a dump records (timeStep time/dt and g_stepNumber).
struct DumpStep {
  double t, dt;
  int nstep;
};

// WriteData(DBL_OUTPUT) + RestartDump: data holds the dumped variables'
// values back to back in file order, names their dbl.out names.
int dbl_write(const char *dir, int nfile, const DumpStep &st, const std::vector<std::string> &names,
              const double *data, size_t count);

// RestartFromFile's reads: record nfile of restart.out into st, dump nfile
// (which must hold exactly `names`, `count` doubles) into data, byte-swapped
// when dbl.out marks the other endianness.
int dbl_read(const char *dir, int nfile, const std::vector<std::string> &names, size_t count,
             DumpStep *st, std::vector<double> &data);

}  // namespace gpl
