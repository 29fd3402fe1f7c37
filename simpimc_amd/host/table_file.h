// table_file.h -- reader of pair-action table files ("spimc-table 1").
//
// A table file holds named datasets whose names are the HDF5 paths the
// reference loaders read (v/diag/r, u_kj_2/data, long_range/u_k, ...).  Each
// dataset is a header line `<name> <f64|i64|str> <count>` followed by values.
#pragma once

#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace spimc_host {

class TableFile {
   public:
    void Load(const std::string& path) {
        std::ifstream f(path);
        if (!f) throw std::runtime_error("cannot open pair-action file " + path);
        std::string magic;
        int version = 0;
        f >> magic >> version;
        if (magic != "spimc-table" || version != 1) throw std::runtime_error(path + ": not a spimc-table 1 file");
        std::string name, kind;
        size_t count = 0;
        while (f >> name >> kind >> count) {
            if (kind == "str") {
                std::string v;
                f >> v;
                strings_[name] = v;
            } else if (kind == "f64" || kind == "i64") {
                std::vector<double> v(count);
                for (size_t i = 0; i < count; ++i)
                    if (!(f >> v[i])) throw std::runtime_error(path + ": short dataset " + name);
                data_[name] = std::move(v);
            } else {
                throw std::runtime_error(path + ": unknown dataset kind " + kind);
            }
        }
        file_ = path;
    }

    const std::vector<double>& Read(const std::string& name) const {
        auto it = data_.find(name);
        if (it == data_.end()) throw std::runtime_error(file_ + ": missing dataset " + name);
        return it->second;
    }

    double Scalar(const std::string& name) const { return Read(name).at(0); }

    std::string String(const std::string& name) const {
        auto it = strings_.find(name);
        if (it == strings_.end()) throw std::runtime_error(file_ + ": missing dataset " + name);
        return it->second;
    }

    bool Has(const std::string& name) const { return data_.count(name) || strings_.count(name); }

   private:
    std::string file_;
    std::map<std::string, std::vector<double>> data_;
    std::map<std::string, std::string> strings_;
};

}  // namespace spimc_host
