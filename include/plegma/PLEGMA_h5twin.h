// PLEGMA_h5twin.h -- the correlator file of a build without HDF5.
//
// The reference writes its correlators with parallel HDF5
// (include/io/PLEGMA_hdf5.h:515-693 driven by lib/PLEGMA_Correlator.cu:809-946).
// This image has no HDF5 library, so the same objects go into a flat record
// file: one record per dataset, holding the HDF5 object path, the element
// type, the shape, the string attributes and the elements in C order.  A
// record file converts to an HDF5 file object for object (plegma_amd.h5twin
// does it with h5py where h5py is installed).
//
//   file   := "PLGH5TW1" record*
//   record := u32 0x54455344 ("DSET")
//             u32 path_len, path            absolute, '/'-separated, no trailing '/'
//             u32 dtype                     1 = f64, 2 = f32, 3 = i32
//             u32 ndim, u64 dims[ndim]
//             u32 nattr, nattr x (u32 name_len, name, u32 value_len, value)
//             u64 nbytes, nbytes of elements, little-endian, C order
//
// All integers are little-endian.  The rules of the reference's writer are
// kept: the file is created when absent and opened for appending otherwise
// (PLEGMA_hdf5.h:570-577); a dataset whose path is already in the file is
// skipped with a warning and the file keeps the old one (PLEGMA_hdf5.h:665-667);
// a path through an existing dataset is an error, as a link that is not a
// group is there (PLEGMA_hdf5.h:272-279).  Header-only, no GPU code.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace plegma {

enum H5TWIN_TYPE : uint32_t { H5TWIN_F64 = 1, H5TWIN_F32 = 2, H5TWIN_I32 = 3 };

template <typename T> struct h5twin_type;
template <> struct h5twin_type<double> { static constexpr uint32_t value = H5TWIN_F64; };
template <> struct h5twin_type<float> { static constexpr uint32_t value = H5TWIN_F32; };
template <> struct h5twin_type<int32_t> { static constexpr uint32_t value = H5TWIN_I32; };

class H5Twin {
public:
  using Attrs = std::vector<std::pair<std::string, std::string>>;
  static constexpr char MAGIC[9] = "PLGH5TW1";
  static constexpr uint32_t TAG = 0x54455344u;

  // "//a//b/" -> "/a/b"
  static std::string normalize(const std::string &p) {
    std::string out = "/";
    for (char ch : p) {
      if (ch == '/' && out.back() == '/') continue;
      out.push_back(ch);
    }
    if (out.size() > 1 && out.back() == '/') out.pop_back();
    return out;
  }

  // create `fname`, or open it and index the datasets it already holds
  explicit H5Twin(const std::string &fname) : fname_(fname) {
    const uint16_t one = 1;
    if (*(const uint8_t *)&one != 1) throw std::runtime_error("the record file is little-endian; this host is not");
    f_ = std::fopen(fname.c_str(), "r+b");
    if (!f_) {
      f_ = std::fopen(fname.c_str(), "w+b");
      if (!f_) throw std::runtime_error("cannot create " + fname);
      put(MAGIC, 8);
      return;
    }
    char magic[8];
    if (std::fread(magic, 1, 8, f_) != 8 || std::memcmp(magic, MAGIC, 8) != 0) {
      std::fclose(f_);
      throw std::runtime_error(fname + " is not a PLGH5TW1 record file");
    }
    uint32_t tag;
    while (std::fread(&tag, 4, 1, f_) == 1) {
      if (tag != TAG) corrupt();
      const std::string path = get_str();
      uint32_t dtype, ndim, nattr;
      get(&dtype, 4);
      get(&ndim, 4);
      if (ndim > 32) corrupt();
      std::vector<uint64_t> dims(ndim);
      if (ndim) get(dims.data(), 8 * ndim);
      get(&nattr, 4);
      for (uint32_t a = 0; a < nattr; a++) {
        get_str();
        get_str();
      }
      uint64_t nbytes;
      get(&nbytes, 8);
      if (std::fseek(f_, (long)nbytes, SEEK_CUR) != 0) corrupt();
      paths_.insert(path);
    }
    if (!std::feof(f_)) corrupt();
    std::clearerr(f_);
  }
  H5Twin(const H5Twin &) = delete;
  H5Twin &operator=(const H5Twin &) = delete;
  ~H5Twin() {
    if (f_) std::fclose(f_);
  }

  bool exists(const std::string &path) const { return paths_.count(normalize(path)) != 0; }

  // Appends the dataset unless `path` is already present.  Returns false
  // (after the reference's warning) when it was skipped.
  template <typename T>
  bool write_dataset(const std::string &path, const T *buf, const std::vector<uint64_t> &dims,
                     const Attrs &attrs = {}) {
    const std::string p = normalize(path);
    if (p == "/") throw std::runtime_error("dataset path is empty");
    if (paths_.count(p)) {
      const size_t cut = p.rfind('/');
      std::fprintf(stderr, "PLEGMA WARNING: An object with name %s already exists in %s. Skipping...\n",
                   p.substr(cut + 1).c_str(), cut ? p.substr(0, cut).c_str() : "/");
      return false;
    }
    for (size_t k = p.find('/', 1); k != std::string::npos; k = p.find('/', k + 1))
      if (paths_.count(p.substr(0, k)))
        throw std::runtime_error("a link with dir " + p.substr(0, k) + " exists but it is not a group");
    const std::string under = p + "/";
    for (const auto &q : paths_)
      if (q.compare(0, under.size(), under) == 0)
        throw std::runtime_error("the group " + p + " exists; it cannot become a dataset");
    uint64_t n = 1;
    for (uint64_t d : dims) n *= d;
    const uint64_t nbytes = n * sizeof(T);
    if (std::fseek(f_, 0, SEEK_END) != 0) throw std::runtime_error("seek failed on " + fname_);
    const uint32_t dtype = h5twin_type<T>::value, ndim = (uint32_t)dims.size(), nattr = (uint32_t)attrs.size();
    put(&TAG, 4);
    put_str(p);
    put(&dtype, 4);
    put(&ndim, 4);
    if (ndim) put(dims.data(), 8 * ndim);
    put(&nattr, 4);
    for (const auto &a : attrs) {
      put_str(a.first);
      put_str(a.second);
    }
    put(&nbytes, 8);
    if (nbytes) put(buf, nbytes);
    if (std::fflush(f_) != 0) throw std::runtime_error("write failed on " + fname_);
    paths_.insert(p);
    return true;
  }

private:
  std::string fname_;
  FILE *f_ = nullptr;
  std::set<std::string> paths_;

  [[noreturn]] void corrupt() { throw std::runtime_error(fname_ + ": truncated or corrupt record"); }
  void get(void *p, size_t n) {
    if (std::fread(p, 1, n, f_) != n) corrupt();
  }
  std::string get_str() {
    uint32_t n;
    get(&n, 4);
    if (n > (1u << 20)) corrupt();
    std::string s(n, '\0');
    if (n) get(&s[0], n);
    return s;
  }
  void put(const void *p, size_t n) {
    if (std::fwrite(p, 1, n, f_) != n) throw std::runtime_error("write failed on " + fname_);
  }
  void put_str(const std::string &s) {
    const uint32_t n = (uint32_t)s.size();
    put(&n, 4);
    put(s.data(), n);
  }
};

}  // namespace plegma
