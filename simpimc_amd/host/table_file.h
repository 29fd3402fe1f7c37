// table_file.h -- reader of pair-action table files.
//
// Two containers hold the same named datasets, whose names are the HDF5 paths
// the reference loaders read (v/diag/r, u_kj_2/data, long_range/u_k, ...):
//  * an HDF5 file (the reference's own format, read through h5_file.h);
//  * the text container "spimc-table 1": per dataset a header line
//    `<name> <f64|i64|str> <count>` followed by the values.
// Multi-dimensional data are flat in storage order; Cube() returns a David
// u_kj cube with the first index (the k,j component) fastest, as the reference
// holds it in memory (david_pair_action_class.h:257-258).
#pragma once

#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "h5_file.h"

namespace spimc_host {

class TableFile {
   public:
    void Load(const std::string& path) {
        file_ = path;
        if (IsHdf5File(path)) {
            t_ = ReadHdf5(path);
            return;
        }
        std::ifstream f(path);
        if (!f) throw std::runtime_error("cannot open pair-action file " + path);
        std::string magic;
        int version = 0;
        f >> magic >> version;
        if (magic != "spimc-table" || version != 1) throw std::runtime_error(path + ": neither HDF5 nor a spimc-table 1 file");
        std::string name, kind;
        size_t count = 0;
        while (f >> name >> kind >> count) {
            if (kind == "str") {
                std::string v;
                f >> v;
                t_.strings[name] = v;
            } else if (kind == "f64" || kind == "i64") {
                std::vector<double> v(count);
                for (size_t i = 0; i < count; ++i)
                    if (!(f >> v[i])) throw std::runtime_error(path + ": short dataset " + name);
                t_.data[name] = std::move(v);
                t_.dims[name] = {count};
                if (kind == "i64") t_.ints.insert(name);
            } else {
                throw std::runtime_error(path + ": unknown dataset kind " + kind);
            }
        }
    }

    const std::vector<double>& Read(const std::string& name) const {
        auto it = t_.data.find(name);
        if (it == t_.data.end()) throw std::runtime_error(file_ + ": missing dataset " + name);
        return it->second;
    }

    double Scalar(const std::string& name) const {
        const auto& v = Read(name);
        if (v.empty()) throw std::runtime_error(file_ + ": empty dataset " + name);
        return v[0];
    }

    std::string String(const std::string& name) const {
        auto it = t_.strings.find(name);
        if (it == t_.strings.end()) throw std::runtime_error(file_ + ": missing dataset " + name);
        return it->second;
    }

    bool Has(const std::string& name) const { return t_.data.count(name) || t_.strings.count(name); }

    // The (n_val, n_grid, n_tau) cube `name`, first index fastest.  A 3-D
    // dataset stored [n_tau][n_grid][n_val] is already in that order; one
    // stored [n_val][n_grid][n_tau] is transposed (when n_val == n_tau the two
    // are indistinguishable and the first is taken); any other shape must hold
    // n_val * n_grid * n_tau values in that order.
    std::vector<double> Cube(const std::string& name, size_t n_val, size_t n_grid, size_t n_tau) const {
        const auto& v = Read(name);
        if (v.size() != n_val * n_grid * n_tau)
            throw std::runtime_error(file_ + ": " + name + " holds " + std::to_string(v.size()) + " values, expected " +
                                     std::to_string(n_val) + " x " + std::to_string(n_grid) + " x " + std::to_string(n_tau));
        const auto& d = t_.dims.at(name);
        const bool tau_fastest = d.size() == 3 && d[0] == n_val && d[1] == n_grid && d[2] == n_tau &&
                                 !(d[0] == n_tau && d[2] == n_val);
        if (!tau_fastest) return v;
        std::vector<double> out(v.size());
        for (size_t c = 0; c < n_val; ++c)
            for (size_t g = 0; g < n_grid; ++g)
                for (size_t t = 0; t < n_tau; ++t) out[c + n_val * (g + n_grid * t)] = v[(c * n_grid + g) * n_tau + t];
        return out;
    }

    // The (n_x, n_y) matrix `name` with x fastest, the memory order of the
    // reference's mat(n_x, n_y) (ilkka_pair_action_class.h:269-272).  A 2-D dataset
    // stored [n_y][n_x] is already in that order; one stored [n_x][n_y] is
    // transposed (when n_x == n_y the first is taken).
    std::vector<double> Matrix(const std::string& name, size_t n_x, size_t n_y) const {
        const auto& v = Read(name);
        if (v.size() != n_x * n_y)
            throw std::runtime_error(file_ + ": " + name + " holds " + std::to_string(v.size()) + " values, expected " +
                                     std::to_string(n_x) + " x " + std::to_string(n_y));
        const auto& d = t_.dims.at(name);
        const bool y_fastest = d.size() == 2 && d[0] == n_x && d[1] == n_y && n_x != n_y;
        if (!y_fastest) return v;
        std::vector<double> out(v.size());
        for (size_t i = 0; i < n_x; ++i)
            for (size_t j = 0; j < n_y; ++j) out[i + n_x * j] = v[i * n_y + j];
        return out;
    }

    const TableData& Data() const { return t_; }

   private:
    std::string file_;
    TableData t_;
};

}  // namespace spimc_host
