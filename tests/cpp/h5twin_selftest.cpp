// h5twin_selftest.cpp -- host-only exercise of the correlator record file
// (include/plegma/PLEGMA_h5twin.h) and of the file-format deduction of
// PLEGMA_Correlator::writeFile: writes known arrays, reopens the file to
// append, hits the skip and error rules, and prints one line per check for
// the Python side, which reads the file back.
//
//   h5twin_selftest <file>
#include <cstdio>
#include <vector>

#include "plegma/PLEGMA_Correlator.h"

using namespace plegma;

int main(int argc, char **argv) {
  if (argc != 2) return 2;
  const std::string fn = argv[1];
  {
    H5Twin w(fn);
    std::vector<double> a(6);
    for (int i = 0; i < 6; i++) a[i] = 0.5 * i - 1.0;
    std::vector<float> b = {1.5f, -2.25f, 3.0f, 1e-30f};
    std::vector<int32_t> m = {0, 0, 0, 1, -1, 0};
    w.write_dataset("/sx01sy02sz03st05/g/x", a.data(), {2, 3}, {{"description", "shape: /a/b"}, {"k", "v"}});
    w.write_dataset("//sx01sy02sz03st05//g/y/", b.data(), {4});
    w.write_dataset("/sx01sy02sz03st05/g/mvec", m.data(), {2, 3});
    std::vector<double> z(6, 9.0);
    std::printf("dup %d\n", (int)w.write_dataset("/sx01sy02sz03st05/g/x", z.data(), {2, 3}));
  }
  {
    H5Twin w(fn);  // reopen: the index holds the three datasets
    std::printf("exists %d %d\n", (int)w.exists("/sx01sy02sz03st05/g/y"), (int)w.exists("/sx01sy02sz03st05/g"));
    std::vector<int32_t> c = {7, 8, 9, 10};
    w.write_dataset("/sx01sy02sz03st05/h/c", c.data(), {2, 2});
    w.write_dataset("/empty", c.data(), {0, 2});
    int errs = 0;
    try {
      w.write_dataset("/sx01sy02sz03st05/g/x/under", c.data(), {1});
    } catch (const std::runtime_error &) {
      errs++;
    }
    try {
      w.write_dataset("/sx01sy02sz03st05/g", c.data(), {1});
    } catch (const std::runtime_error &) {
      errs++;
    }
    std::printf("errors %d\n", errs);
  }
  {
    std::FILE *f = std::fopen((fn + ".bad").c_str(), "wb");
    std::fputs("NOTATWIN", f);
    std::fclose(f);
    int bad = 0;
    try {
      H5Twin w(fn + ".bad");
    } catch (const std::runtime_error &) {
      bad = 1;
    }
    std::printf("bad %d\n", bad);
  }
  using C = PLEGMA_Correlator<double>;
  std::printf("formats %d %d %d %d %d %d %d\n", (int)C::deduceFormat("a/b.h5"), (int)C::deduceFormat("b.hdf5"),
              (int)C::deduceFormat("b.h5/grp/sub"), (int)C::deduceFormat("twop.dat"),
              (int)C::deduceFormat("x.txt"), (int)C::deduceFormat("x.lime"), (int)C::deduceFormat("plain"));
  return 0;
}
