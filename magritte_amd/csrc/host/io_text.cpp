// io_text.cpp -- IoText: the text back end of the io model layout
// (behaviour of src/Io/cpp/io_cpp_text.cpp: one file per dataset, 16
// significant digits in scientific notation, tab-separated array rows).
#include <sys/stat.h>

#include <fstream>
#include <iomanip>
#include <sstream>

#include "io.hpp"

namespace magritte {

namespace {

bool path_exists(const std::string &path)
{
    struct stat buffer;
    return stat(path.c_str(), &buffer) == 0;
}

// Creates the directories of "<io_file><name>" so that a write succeeds.
void make_parents(const std::string &path)
{
    for (size_t pos = path.find('/', 1); pos != std::string::npos; pos = path.find('/', pos + 1)) {
        const std::string dir = path.substr(0, pos);
        if (!path_exists(dir)) mkdir(dir.c_str(), 0755);
    }
}

template <typename T>
int read_one(const std::string &fname, T &value)
{
    std::ifstream file(fname);
    if (!file.is_open()) return -1;
    file >> value;
    return file.fail() ? -1 : 0;
}

template <typename T>
int write_one(const std::string &fname, const T &value)
{
    make_parents(fname);
    std::ofstream file(fname);
    if (!file.is_open()) return -1;
    file << std::scientific << std::setprecision(16) << value;
    return 0;
}

template <typename T>
int read_list_t(const std::string &fname, std::vector<T> &list)
{
    std::ifstream file(fname);
    if (!file.is_open()) return -1;
    size_t n = 0;
    T v;
    while (file >> v) {
        if (n < list.size()) list[n] = v;
        else list.push_back(v);
        n++;
    }
    return 0;
}

template <typename T>
int write_list_t(const std::string &fname, const std::vector<T> &list)
{
    make_parents(fname);
    std::ofstream file(fname);
    if (!file.is_open()) return -1;
    file << std::scientific << std::setprecision(16);
    for (const T &v : list) file << v << '\n';
    return 0;
}

template <typename T>
int read_array_t(const std::string &fname, std::vector<std::vector<T>> &array)
{
    std::ifstream file(fname);
    if (!file.is_open()) return -1;
    std::string line;
    for (size_t n1 = 0; n1 < array.size(); n1++) {
        if (!std::getline(file, line)) return -1;
        std::stringstream ss(line);
        for (size_t n2 = 0; n2 < array[n1].size(); n2++) ss >> array[n1][n2];
    }
    return 0;
}

template <typename T>
int write_array_t(const std::string &fname, const std::vector<std::vector<T>> &array)
{
    make_parents(fname);
    std::ofstream file(fname);
    if (!file.is_open()) return -1;
    file << std::scientific << std::setprecision(16);
    for (const auto &row : array) {
        for (const T &v : row) file << v << '\t';
        file << '\n';
    }
    return 0;
}

}  // namespace

int IoText::read_length(const std::string &name, size_t &length) const
{
    const std::string fname = io_file + name;
    length = 0;
    if (path_exists(fname + ".txt")) {
        std::ifstream file(fname + ".txt");
        std::string line;
        while (std::getline(file, line)) length++;
    } else {
        while (path_exists(fname + std::to_string(length))) length++;
    }
    return 0;
}

int IoText::read_width(const std::string &name, size_t &width) const
{
    const std::string fname = io_file + name;
    width = 0;
    if (path_exists(fname + ".txt")) {
        std::ifstream file(fname + ".txt");
        std::string line, elem;
        std::getline(file, line);
        std::stringstream ss(line);
        while (std::getline(ss, elem, '\t')) width++;
    } else {
        while (path_exists(fname + std::to_string(width))) width++;
    }
    return 0;
}

int IoText::read_number(const std::string &name, size_t &number) const { return read_one(io_file + name + ".txt", number); }
int IoText::read_number(const std::string &name, long &number) const { return read_one(io_file + name + ".txt", number); }
int IoText::read_number(const std::string &name, double &number) const { return read_one(io_file + name + ".txt", number); }
int IoText::read_word(const std::string &name, std::string &word) const { return read_one(io_file + name + ".txt", word); }

int IoText::read_bool(const std::string &name, bool &value) const
{
    std::string word;
    int err = read_word(name, word);
    if (word == "true") value = true;
    else if (word == "false") value = false;
    else err = -1;
    return err;
}

int IoText::read_list(const std::string &name, Long1 &list) const { return read_list_t(io_file + name + ".txt", list); }
int IoText::read_list(const std::string &name, Double1 &list) const { return read_list_t(io_file + name + ".txt", list); }
int IoText::read_list(const std::string &name, String1 &list) const { return read_list_t(io_file + name + ".txt", list); }
int IoText::read_array(const std::string &name, Long2 &array) const { return read_array_t(io_file + name + ".txt", array); }
int IoText::read_array(const std::string &name, Double2 &array) const { return read_array_t(io_file + name + ".txt", array); }

int IoText::write_number(const std::string &name, const size_t &number) const { return write_one(io_file + name + ".txt", number); }
int IoText::write_number(const std::string &name, const long &number) const { return write_one(io_file + name + ".txt", number); }
int IoText::write_number(const std::string &name, const double &number) const { return write_one(io_file + name + ".txt", number); }
int IoText::write_word(const std::string &name, const std::string &word) const { return write_one(io_file + name + ".txt", word); }
int IoText::write_bool(const std::string &name, const bool &value) const { return write_word(name, value ? "true" : "false"); }
int IoText::write_list(const std::string &name, const Long1 &list) const { return write_list_t(io_file + name + ".txt", list); }
int IoText::write_list(const std::string &name, const Double1 &list) const { return write_list_t(io_file + name + ".txt", list); }
int IoText::write_list(const std::string &name, const String1 &list) const { return write_list_t(io_file + name + ".txt", list); }
int IoText::write_array(const std::string &name, const Long2 &array) const { return write_array_t(io_file + name + ".txt", array); }
int IoText::write_array(const std::string &name, const Double2 &array) const { return write_array_t(io_file + name + ".txt", array); }

}  // namespace magritte
