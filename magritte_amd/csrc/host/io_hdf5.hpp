// io_hdf5.hpp -- IoHDF5, the HDF5 back end of the io model layout
// (the reference's IoPython("hdf5") over src/Io/python/io_python_hdf5.py),
// built when the HDF5 C library is available.
#pragma once

#include "io.hpp"

namespace magritte {

struct IoHDF5 : public Io {
    explicit IoHDF5(const std::string &file) : Io(file) {}

    int read_length(const std::string &name, size_t &length) const override;
    int read_width(const std::string &name, size_t &width) const override;
    int read_number(const std::string &name, size_t &number) const override;
    int read_number(const std::string &name, long &number) const override;
    int read_number(const std::string &name, double &number) const override;
    int read_word(const std::string &name, std::string &word) const override;
    int read_bool(const std::string &name, bool &value) const override;
    int read_list(const std::string &name, Long1 &list) const override;
    int read_list(const std::string &name, Double1 &list) const override;
    int read_list(const std::string &name, String1 &list) const override;
    int read_array(const std::string &name, Long2 &array) const override;
    int read_array(const std::string &name, Double2 &array) const override;

    int write_number(const std::string &name, const size_t &number) const override;
    int write_number(const std::string &name, const long &number) const override;
    int write_number(const std::string &name, const double &number) const override;
    int write_word(const std::string &name, const std::string &word) const override;
    int write_bool(const std::string &name, const bool &value) const override;
    int write_list(const std::string &name, const Long1 &list) const override;
    int write_list(const std::string &name, const Double1 &list) const override;
    int write_list(const std::string &name, const String1 &list) const override;
    int write_array(const std::string &name, const Long2 &array) const override;
    int write_array(const std::string &name, const Double2 &array) const override;
};

// IoPython (src/Io/python/io_python.hpp): the reference selects the back end
// by name, IoPython("hdf5", file); here the HDF5 one is native C++.
struct IoPython : public IoHDF5 {
    std::string implementation;
    IoPython(const std::string &imp, const std::string &file);
};

}  // namespace magritte
