// h5_file.h -- the HDF5 side of the host driver: reading pair-action tables
// written as HDF5 files and writing the block output in the reference's layout.
//
// libhdf5 is loaded at run time (dlopen) on the first HDF5 file, so the driver
// links against nothing beyond the C ABI library and a run with text tables
// needs no HDF5 at all.  Prototypes, types and enumerations come from the
// library's own header (found at build time, SPIMC_HAVE_HDF5); the library
// found at run time must have the header's major.minor version.  Search order:
// $SPIMC_HDF5_LIB, the library directory that sat beside the header at build
// time (SPIMC_HDF5_LIB_DIR), then the loader's default search for libhdf5.so.
//
// Objects and names follow the reference's output file
// (framework/framework_class.h:68-80, framework/simulation_class.h:18-29,
// events/moves/move_class.h:49-53,80-94, events/observables/observable_class.h:84-87):
// scalars and strings are one-element 1-D datasets, a per-block scalar is an
// extendable 1-D dataset and a per-block histogram an extendable 2-D dataset
// [block][bin].
#pragma once

#include <dlfcn.h>

#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#if SPIMC_HAVE_HDF5
#include <hdf5.h>
#endif

namespace spimc_host {

inline bool IsHdf5File(const std::string& path) {
    static const unsigned char sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    unsigned char b[8] = {0};
    const size_t n = std::fread(b, 1, 8, f);
    std::fclose(f);
    return n == 8 && std::memcmp(b, sig, 8) == 0;
}

// A table file's contents: numeric datasets as flat double vectors in the
// file's storage order, with their dimensions, and string datasets by their
// first element.
struct TableData {
    std::map<std::string, std::vector<double>> data;
    std::map<std::string, std::vector<uint64_t>> dims;
    std::set<std::string> ints;
    std::map<std::string, std::string> strings;
};

#if SPIMC_HAVE_HDF5

class H5Api {
   public:
    static H5Api& Get() {
        static H5Api api;
        if (!api.so_) throw std::runtime_error("HDF5 library not found: " + api.why_);
        return api;
    }

#define SPIMC_H5_FN(name) decltype(&::name) name = nullptr
    SPIMC_H5_FN(H5open);
    SPIMC_H5_FN(H5get_libversion);
    SPIMC_H5_FN(H5free_memory);
    SPIMC_H5_FN(H5Eset_auto2);
    SPIMC_H5_FN(H5Fopen);
    SPIMC_H5_FN(H5Fcreate);
    SPIMC_H5_FN(H5Fflush);
    SPIMC_H5_FN(H5Fclose);
    SPIMC_H5_FN(H5Lvisit);
    SPIMC_H5_FN(H5Lexists);
    SPIMC_H5_FN(H5Oopen);
    SPIMC_H5_FN(H5Oclose);
    SPIMC_H5_FN(H5Iget_type);
    SPIMC_H5_FN(H5Gcreate2);
    SPIMC_H5_FN(H5Gclose);
    SPIMC_H5_FN(H5Dopen2);
    SPIMC_H5_FN(H5Dcreate2);
    SPIMC_H5_FN(H5Dget_type);
    SPIMC_H5_FN(H5Dget_space);
    SPIMC_H5_FN(H5Dread);
    SPIMC_H5_FN(H5Dwrite);
    SPIMC_H5_FN(H5Dset_extent);
    SPIMC_H5_FN(H5Dclose);
    SPIMC_H5_FN(H5Screate_simple);
    SPIMC_H5_FN(H5Sget_simple_extent_ndims);
    SPIMC_H5_FN(H5Sget_simple_extent_dims);
    SPIMC_H5_FN(H5Sget_simple_extent_npoints);
    SPIMC_H5_FN(H5Sselect_hyperslab);
    SPIMC_H5_FN(H5Sclose);
    SPIMC_H5_FN(H5Tget_class);
    SPIMC_H5_FN(H5Tget_size);
    SPIMC_H5_FN(H5Tis_variable_str);
    SPIMC_H5_FN(H5Tcopy);
    SPIMC_H5_FN(H5Tset_size);
    SPIMC_H5_FN(H5Tclose);
    SPIMC_H5_FN(H5Pcreate);
    SPIMC_H5_FN(H5Pset_chunk);
    SPIMC_H5_FN(H5Pclose);
#undef SPIMC_H5_FN
    // predefined type and property-class ids (the library's *_g globals, valid after H5open)
    hid_t t_double = -1, t_int64 = -1, t_c_s1 = -1, p_dataset_create = -1;

    // flags of H5Fopen / H5Fcreate (H5F_ACC_RDONLY, H5F_ACC_TRUNC)
    static constexpr unsigned kAccRdonly = 0x0000u, kAccTrunc = 0x0002u;

    void Check(long long rc, const std::string& what) const {
        if (rc < 0) throw std::runtime_error("HDF5: " + what);
    }

   private:
    void* so_ = nullptr;
    std::string why_;

    H5Api() {
        std::vector<std::string> tries;
        if (const char* e = std::getenv("SPIMC_HDF5_LIB")) tries.push_back(e);
#ifdef SPIMC_HDF5_LIB_DIR
        tries.push_back(std::string(SPIMC_HDF5_LIB_DIR) + "/libhdf5.so");
#endif
        tries.push_back("libhdf5.so");
        for (auto& t : tries) {
            so_ = dlopen(t.c_str(), RTLD_NOW | RTLD_LOCAL);
            if (so_) break;
            why_ += (why_.empty() ? "" : "; ") + std::string(dlerror());
        }
        if (!so_) return;
        try {
            Resolve();
        } catch (const std::exception& ex) {
            dlclose(so_);
            so_ = nullptr;
            why_ = ex.what();
        }
    }

    template <class F>
    void Sym(F& f, const char* name) {
        f = reinterpret_cast<F>(dlsym(so_, name));
        if (!f) throw std::runtime_error(std::string("missing symbol ") + name);
    }
    hid_t Global(const char* name) {
        auto* p = reinterpret_cast<const hid_t*>(dlsym(so_, name));
        if (!p) throw std::runtime_error(std::string("missing symbol ") + name);
        return *p;
    }

    void Resolve() {
#define SPIMC_H5_SYM(name) Sym(name, #name)
        SPIMC_H5_SYM(H5open);
        SPIMC_H5_SYM(H5get_libversion);
        SPIMC_H5_SYM(H5free_memory);
        SPIMC_H5_SYM(H5Eset_auto2);
        SPIMC_H5_SYM(H5Fopen);
        SPIMC_H5_SYM(H5Fcreate);
        SPIMC_H5_SYM(H5Fflush);
        SPIMC_H5_SYM(H5Fclose);
        SPIMC_H5_SYM(H5Lvisit);
        SPIMC_H5_SYM(H5Lexists);
        SPIMC_H5_SYM(H5Oopen);
        SPIMC_H5_SYM(H5Oclose);
        SPIMC_H5_SYM(H5Iget_type);
        SPIMC_H5_SYM(H5Gcreate2);
        SPIMC_H5_SYM(H5Gclose);
        SPIMC_H5_SYM(H5Dopen2);
        SPIMC_H5_SYM(H5Dcreate2);
        SPIMC_H5_SYM(H5Dget_type);
        SPIMC_H5_SYM(H5Dget_space);
        SPIMC_H5_SYM(H5Dread);
        SPIMC_H5_SYM(H5Dwrite);
        SPIMC_H5_SYM(H5Dset_extent);
        SPIMC_H5_SYM(H5Dclose);
        SPIMC_H5_SYM(H5Screate_simple);
        SPIMC_H5_SYM(H5Sget_simple_extent_ndims);
        SPIMC_H5_SYM(H5Sget_simple_extent_dims);
        SPIMC_H5_SYM(H5Sget_simple_extent_npoints);
        SPIMC_H5_SYM(H5Sselect_hyperslab);
        SPIMC_H5_SYM(H5Sclose);
        SPIMC_H5_SYM(H5Tget_class);
        SPIMC_H5_SYM(H5Tget_size);
        SPIMC_H5_SYM(H5Tis_variable_str);
        SPIMC_H5_SYM(H5Tcopy);
        SPIMC_H5_SYM(H5Tset_size);
        SPIMC_H5_SYM(H5Tclose);
        SPIMC_H5_SYM(H5Pcreate);
        SPIMC_H5_SYM(H5Pset_chunk);
        SPIMC_H5_SYM(H5Pclose);
#undef SPIMC_H5_SYM
        Check(H5open(), "H5open");
        unsigned maj = 0, min = 0, rel = 0;
        Check(H5get_libversion(&maj, &min, &rel), "H5get_libversion");
        if (maj != H5_VERS_MAJOR || min != H5_VERS_MINOR)
            throw std::runtime_error("libhdf5 " + std::to_string(maj) + "." + std::to_string(min) + " found, built against " +
                                     std::to_string(H5_VERS_MAJOR) + "." + std::to_string(H5_VERS_MINOR));
        t_double = Global("H5T_NATIVE_DOUBLE_g");
        t_int64 = Global("H5T_NATIVE_LLONG_g");
        t_c_s1 = Global("H5T_C_S1_g");
        p_dataset_create = Global("H5P_CLS_DATASET_CREATE_ID_g");
        // errors are reported by return codes, not printed by the library
        Check(H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr), "H5Eset_auto2");
    }
};

// Every dataset of an HDF5 file, by its path without the leading '/'.
inline TableData ReadHdf5(const std::string& path) {
    H5Api& h = H5Api::Get();
    const hid_t f = h.H5Fopen(path.c_str(), H5Api::kAccRdonly, H5P_DEFAULT);
    if (f < 0) throw std::runtime_error("cannot open HDF5 file " + path);
    std::vector<std::string> names;
    auto visit = [](hid_t, const char* name, const H5L_info_t*, void* op) -> herr_t {
        static_cast<std::vector<std::string>*>(op)->push_back(name);
        return 0;
    };
    TableData t;
    try {
        h.Check(h.H5Lvisit(f, H5_INDEX_NAME, H5_ITER_NATIVE, visit, &names), path + ": link traversal");
        for (const auto& name : names) {
            const hid_t o = h.H5Oopen(f, name.c_str(), H5P_DEFAULT);
            if (o < 0) continue;
            const bool is_ds = h.H5Iget_type(o) == H5I_DATASET;
            h.H5Oclose(o);
            if (!is_ds) continue;
            const hid_t ds = h.H5Dopen2(f, name.c_str(), H5P_DEFAULT);
            const hid_t ty = h.H5Dget_type(ds), sp = h.H5Dget_space(ds);
            const int nd = h.H5Sget_simple_extent_ndims(sp);
            std::vector<hsize_t> d(nd > 0 ? nd : 0);
            if (nd > 0) h.H5Sget_simple_extent_dims(sp, d.data(), nullptr);
            const hssize_t np = h.H5Sget_simple_extent_npoints(sp);
            const H5T_class_t cls = h.H5Tget_class(ty);
            herr_t rc = 0;
            if (cls == H5T_STRING && np > 0) {
                const hid_t mt = h.H5Tcopy(h.t_c_s1);
                if (h.H5Tis_variable_str(ty) > 0) {
                    h.H5Tset_size(mt, H5T_VARIABLE);
                    std::vector<char*> buf(np, nullptr);
                    rc = h.H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data());
                    if (rc >= 0) t.strings[name] = buf[0] ? buf[0] : "";
                    for (char* s : buf)
                        if (s) h.H5free_memory(s);
                } else {
                    const size_t len = h.H5Tget_size(ty);
                    h.H5Tset_size(mt, len);
                    std::vector<char> buf(len * np + 1, 0);
                    rc = h.H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data());
                    std::string s(buf.data(), len);
                    s = s.substr(0, s.find('\0'));
                    while (!s.empty() && s.back() == ' ') s.pop_back();
                    if (rc >= 0) t.strings[name] = s;
                }
                h.H5Tclose(mt);
            } else if (cls == H5T_INTEGER || cls == H5T_FLOAT) {
                std::vector<double> v(np);
                if (np > 0) rc = h.H5Dread(ds, h.t_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
                t.data[name] = std::move(v);
                t.dims[name].assign(d.begin(), d.end());
                if (cls == H5T_INTEGER) t.ints.insert(name);
            }
            h.H5Sclose(sp);
            h.H5Tclose(ty);
            h.H5Dclose(ds);
            h.Check(rc, path + ": reading " + name);
        }
    } catch (...) {
        h.H5Fclose(f);
        throw;
    }
    h.H5Fclose(f);
    return t;
}

// Writer of one HDF5 output file; groups are created on demand.
class H5Out {
   public:
    explicit H5Out(const std::string& path) : h_(H5Api::Get()), path_(path) {
        f_ = h_.H5Fcreate(path.c_str(), H5Api::kAccTrunc, H5P_DEFAULT, H5P_DEFAULT);
        if (f_ < 0) throw std::runtime_error("cannot create HDF5 file " + path);
    }
    ~H5Out() {
        for (auto& kv : ext_) h_.H5Dclose(kv.second);
        h_.H5Fclose(f_);
    }
    H5Out(const H5Out&) = delete;
    H5Out& operator=(const H5Out&) = delete;

    void Group(const std::string& p) {
        std::string cur;
        size_t pos = 0;
        while (pos < p.size()) {
            size_t next = p.find('/', pos);
            if (next == std::string::npos) next = p.size();
            if (next > pos) {
                cur += (cur.empty() ? "" : "/") + p.substr(pos, next - pos);
                if (h_.H5Lexists(f_, cur.c_str(), H5P_DEFAULT) <= 0) {
                    const hid_t g = h_.H5Gcreate2(f_, cur.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
                    h_.Check(g, path_ + ": group " + cur);
                    h_.H5Gclose(g);
                }
            }
            pos = next + 1;
        }
    }

    void String(const std::string& name, const std::string& s) {
        const hid_t ty = h_.H5Tcopy(h_.t_c_s1);
        h_.H5Tset_size(ty, s.size() + 1);
        Fixed(name, ty, {1}, s.c_str());
        h_.H5Tclose(ty);
    }
    void Double(const std::string& name, double v) { Fixed(name, h_.t_double, {1}, &v); }
    void Int(const std::string& name, long long v) { Fixed(name, h_.t_int64, {1}, &v); }
    void Doubles(const std::string& name, const std::vector<double>& v, std::vector<hsize_t> dims = {}) {
        if (dims.empty()) dims = {v.size()};
        Fixed(name, h_.t_double, dims, v.data());
    }
    void Ints(const std::string& name, const std::vector<long long>& v, std::vector<hsize_t> dims = {}) {
        if (dims.empty()) dims = {v.size()};
        Fixed(name, h_.t_int64, dims, v.data());
    }

    // Appends one block to an extendable dataset: a scalar row (n == 0) of a
    // 1-D dataset, or a row of n values of a 2-D dataset [block][n].
    void Append(const std::string& name, const double* row, size_t n) {
        AppendRow(name, h_.t_double, row, n ? std::vector<hsize_t>{n} : std::vector<hsize_t>{});
    }
    void AppendInt(const std::string& name, long long v) { AppendRow(name, h_.t_int64, &v, {}); }
    void AppendInts(const std::string& name, const std::vector<long long>& row) {
        AppendRow(name, h_.t_int64, row.data(), {(hsize_t)row.size()});
    }
    // Appends one block of shape `row_dims` to the dataset [block][row_dims...].
    void Append(const std::string& name, const double* row, const std::vector<hsize_t>& row_dims) {
        AppendRow(name, h_.t_double, row, row_dims);
    }
    // A one-element integer dataset, created or overwritten.
    void SetInt(const std::string& name, long long v) {
        if (h_.H5Lexists(f_, name.c_str(), H5P_DEFAULT) <= 0) return Int(name, v);
        const hid_t ds = h_.H5Dopen2(f_, name.c_str(), H5P_DEFAULT);
        h_.Check(ds, path_ + ": opening " + name);
        const herr_t rc = h_.H5Dwrite(ds, h_.t_int64, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v);
        h_.H5Dclose(ds);
        h_.Check(rc, path_ + ": rewriting " + name);
    }

    void Flush() { h_.H5Fflush(f_, H5F_SCOPE_LOCAL); }

   private:
    H5Api& h_;
    std::string path_;
    hid_t f_ = -1;
    std::map<std::string, hid_t> ext_;
    std::map<std::string, hsize_t> rows_;

    void Parent(const std::string& name) {
        const size_t s = name.rfind('/');
        if (s != std::string::npos) Group(name.substr(0, s));
    }

    void Fixed(const std::string& name, hid_t ty, const std::vector<hsize_t>& dims, const void* buf) {
        Parent(name);
        const hid_t sp = h_.H5Screate_simple((int)dims.size(), dims.data(), nullptr);
        const hid_t ds = h_.H5Dcreate2(f_, name.c_str(), ty, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        herr_t rc = ds < 0 ? -1 : h_.H5Dwrite(ds, ty, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
        if (ds >= 0) h_.H5Dclose(ds);
        h_.H5Sclose(sp);
        h_.Check(rc, path_ + ": writing " + name);
    }

    void AppendRow(const std::string& name, hid_t ty, const void* row, const std::vector<hsize_t>& row_dims) {
        const int rank = 1 + (int)row_dims.size();
        std::vector<hsize_t> d0{0}, dmax{H5S_UNLIMITED}, chunk{row_dims.empty() ? (hsize_t)256 : (hsize_t)1};
        for (hsize_t d : row_dims) {
            d0.push_back(d);
            dmax.push_back(d);
            chunk.push_back(d);
        }
        auto it = ext_.find(name);
        if (it == ext_.end()) {
            Parent(name);
            const hid_t sp = h_.H5Screate_simple(rank, d0.data(), dmax.data());
            const hid_t pl = h_.H5Pcreate(h_.p_dataset_create);
            h_.H5Pset_chunk(pl, rank, chunk.data());
            const hid_t ds = h_.H5Dcreate2(f_, name.c_str(), ty, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
            h_.H5Pclose(pl);
            h_.H5Sclose(sp);
            h_.Check(ds, path_ + ": creating " + name);
            it = ext_.emplace(name, ds).first;
            rows_[name] = 0;
        }
        const hid_t ds = it->second;
        hsize_t& rows = rows_[name];
        std::vector<hsize_t> size(d0), start(rank, 0), count(d0);
        size[0] = rows + 1;
        start[0] = rows;
        count[0] = 1;
        h_.Check(h_.H5Dset_extent(ds, size.data()), path_ + ": extending " + name);
        const hid_t fs = h_.H5Dget_space(ds);
        h_.H5Sselect_hyperslab(fs, H5S_SELECT_SET, start.data(), nullptr, count.data(), nullptr);
        const hid_t ms = h_.H5Screate_simple(rank, count.data(), nullptr);
        const herr_t rc = h_.H5Dwrite(ds, ty, ms, fs, H5P_DEFAULT, row);
        h_.H5Sclose(ms);
        h_.H5Sclose(fs);
        h_.Check(rc, path_ + ": appending to " + name);
        ++rows;
    }
};

#else  // no HDF5 header at build time

inline TableData ReadHdf5(const std::string& path) {
    throw std::runtime_error(path + " is an HDF5 file, and this driver was built without HDF5 support");
}

#endif

}  // namespace spimc_host
