// io.hpp -- the io model layout back end of the drop-in host layer.
//
// Io is the abstract reader/writer of the reference (src/Io/io.hpp:17-65);
// IoText stores every dataset as "<io_file><name>.txt" (src/Io/cpp/io_cpp_text.cpp):
// a number or word per file, lists one value per line, arrays one
// tab-separated row per line.  read_length counts the lines of a file, or,
// when the file does not exist, the paths "<name>0", "<name>1", ... (groups).
#pragma once

#include <string>
#include <vector>

namespace magritte {

using Double1 = std::vector<double>;
using Double2 = std::vector<Double1>;
using Double3 = std::vector<Double2>;
using Long1   = std::vector<long>;
using Long2   = std::vector<Long1>;
using Long3   = std::vector<Long2>;
using Size1   = std::vector<size_t>;
using Bool1   = std::vector<bool>;
using String1 = std::vector<std::string>;

struct Io {
    std::string io_file;
    explicit Io(const std::string &file) : io_file(file) {}
    virtual ~Io() = default;

    virtual int read_length(const std::string &name, size_t &length) const = 0;
    virtual int read_width(const std::string &name, size_t &width) const = 0;
    virtual int read_number(const std::string &name, size_t &number) const = 0;
    virtual int read_number(const std::string &name, long &number) const = 0;
    virtual int read_number(const std::string &name, double &number) const = 0;
    virtual int read_word(const std::string &name, std::string &word) const = 0;
    virtual int read_bool(const std::string &name, bool &value) const = 0;
    virtual int read_list(const std::string &name, Long1 &list) const = 0;
    virtual int read_list(const std::string &name, Double1 &list) const = 0;
    virtual int read_list(const std::string &name, String1 &list) const = 0;
    virtual int read_array(const std::string &name, Long2 &array) const = 0;
    virtual int read_array(const std::string &name, Double2 &array) const = 0;

    virtual int write_number(const std::string &name, const size_t &number) const = 0;
    virtual int write_number(const std::string &name, const long &number) const = 0;
    virtual int write_number(const std::string &name, const double &number) const = 0;
    virtual int write_word(const std::string &name, const std::string &word) const = 0;
    virtual int write_bool(const std::string &name, const bool &value) const = 0;
    virtual int write_list(const std::string &name, const Long1 &list) const = 0;
    virtual int write_list(const std::string &name, const Double1 &list) const = 0;
    virtual int write_list(const std::string &name, const String1 &list) const = 0;
    virtual int write_array(const std::string &name, const Long2 &array) const = 0;
    virtual int write_array(const std::string &name, const Double2 &array) const = 0;
};

struct IoText : public Io {
    explicit IoText(const std::string &file) : Io(file) {}

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

}  // namespace magritte
