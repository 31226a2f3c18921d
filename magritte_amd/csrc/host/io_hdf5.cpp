// io_hdf5.cpp -- IoHDF5: the reference's HDF5 io layout
// (src/Io/python/io_python_hdf5.py behind IoPython("hdf5"),
// src/Io/python/io_python.cpp) written against the HDF5 C library instead of
// an embedded Python interpreter with h5py.
//
// Layout, as the reference's:
//  * "<object>/.<attr>" names are attributes of the group <object> (the root
//    for ".<attr>"): numbers (64-bit integers or doubles), words and booleans
//    ("true"/"false", as IoPython::write_bool stores them) as UTF-8 strings;
//  * every other name is a dataset at that path: lists 1-D, arrays 2-D;
//    groups on the way are created, an existing dataset is replaced, empty
//    lists and arrays are not written (IoPython::write_list / write_array);
//  * read_length of a dataset is its first dimension, otherwise the number of
//    members of the containing group whose name contains the last component
//    (how the reference counts "LineProducingSpecies_<l>" groups).
// Every call opens and closes the file, as each h5py call does.
#include "io_hdf5.hpp"

#include <hdf5.h>

#include <cstring>
#include <stdexcept>

namespace magritte {

namespace {

// HDF5 prints an error stack on every failed call; reads of absent optional
// datasets are expected, so printing is switched off once.
void quiet()
{
    static bool done = false;
    if (!done) {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        done = true;
    }
}

bool file_exists(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fclose(f);
    return true;
}

// RAII closers for the identifier kinds used here
struct Fid {
    hid_t id = -1;
    ~Fid() { if (id >= 0) H5Fclose(id); }
};
struct Gid {
    hid_t id = -1;
    ~Gid() { if (id >= 0) H5Oclose(id); }
};
struct Did {
    hid_t id = -1;
    ~Did() { if (id >= 0) H5Dclose(id); }
};
struct Aid {
    hid_t id = -1;
    ~Aid() { if (id >= 0) H5Aclose(id); }
};
struct Sid {
    hid_t id = -1;
    ~Sid() { if (id >= 0) H5Sclose(id); }
};
struct Tid {
    hid_t id = -1;
    ~Tid() { if (id >= 0) H5Tclose(id); }
};
struct Pid {
    hid_t id = -1;
    ~Pid() { if (id >= 0) H5Pclose(id); }
};

hid_t open_read(const std::string &path)
{
    quiet();
    if (!file_exists(path)) return -1;
    return H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
}

hid_t open_write(const std::string &path)
{
    quiet();
    if (file_exists(path)) return H5Fopen(path.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
    return H5Fcreate(path.c_str(), H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
}

std::string strip(const std::string &p)
{
    std::string s = p;
    while (s.size() > 1 && s.back() == '/') s.pop_back();
    return s.empty() ? std::string("/") : s;
}

// "<object>/.<attr>" -> (object path, attribute), as name.split('.')[0:2]
bool split_attribute(const std::string &name, std::string &object, std::string &attr)
{
    const size_t dot = name.find('.');
    if (dot == std::string::npos) return false;
    object = strip(name.substr(0, dot));
    const size_t dot2 = name.find('.', dot + 1);
    attr = name.substr(dot + 1, dot2 == std::string::npos ? std::string::npos : dot2 - dot - 1);
    return !attr.empty();
}

bool link_exists(hid_t file, const std::string &path)
{
    // every component must exist for H5Lexists to be defined
    std::string part;
    size_t start = (path[0] == '/') ? 1 : 0;
    while (start <= path.size()) {
        const size_t slash = path.find('/', start);
        part = path.substr(0, slash == std::string::npos ? path.size() : slash);
        if (!part.empty() && part != "/" && H5Lexists(file, part.c_str(), H5P_DEFAULT) <= 0) return false;
        if (slash == std::string::npos) break;
        start = slash + 1;
    }
    return true;
}

hid_t link_create_plist()
{
    const hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
    H5Pset_create_intermediate_group(lcpl, 1);
    return lcpl;
}

// group at `path`, created with its parents when absent
hid_t require_group(hid_t file, const std::string &path)
{
    const std::string p = strip(path);
    if (p == "/") return H5Oopen(file, "/", H5P_DEFAULT);
    if (link_exists(file, p)) return H5Oopen(file, p.c_str(), H5P_DEFAULT);
    Pid lcpl;
    lcpl.id = link_create_plist();
    return H5Gcreate2(file, p.c_str(), lcpl.id, H5P_DEFAULT, H5P_DEFAULT);
}

hid_t vlen_string_type()
{
    const hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, H5T_VARIABLE);
    H5Tset_cset(t, H5T_CSET_UTF8);
    return t;
}

template <typename T> hid_t native();
template <> hid_t native<long>() { return H5T_NATIVE_INT64; }
template <> hid_t native<size_t>() { return H5T_NATIVE_UINT64; }
template <> hid_t native<double>() { return H5T_NATIVE_DOUBLE; }

template <typename T>
int read_attribute_number(const std::string &file_name, const std::string &name, T &value)
{
    std::string obj, attr;
    if (!split_attribute(name, obj, attr)) return -1;
    Fid f;
    if ((f.id = open_read(file_name)) < 0) return -1;
    if (obj != "/" && !link_exists(f.id, obj)) return -1;
    Gid g;
    if ((g.id = H5Oopen(f.id, obj.c_str(), H5P_DEFAULT)) < 0) return -1;
    if (H5Aexists(g.id, attr.c_str()) <= 0) return -1;
    Aid a;
    if ((a.id = H5Aopen(g.id, attr.c_str(), H5P_DEFAULT)) < 0) return -1;
    return H5Aread(a.id, native<T>(), &value) < 0 ? -1 : 0;
}

// writes a scalar attribute of type `type` from `buf`, replacing an existing one
int write_attribute(const std::string &file_name, const std::string &name, hid_t type, const void *buf)
{
    std::string obj, attr;
    if (!split_attribute(name, obj, attr)) return -1;
    Fid f;
    if ((f.id = open_write(file_name)) < 0) return -1;
    Gid g;
    if ((g.id = require_group(f.id, obj)) < 0) return -1;
    if (H5Aexists(g.id, attr.c_str()) > 0 && H5Adelete(g.id, attr.c_str()) < 0) return -1;
    Sid s;
    s.id = H5Screate(H5S_SCALAR);
    Aid a;
    if ((a.id = H5Acreate2(g.id, attr.c_str(), type, s.id, H5P_DEFAULT, H5P_DEFAULT)) < 0) return -1;
    return H5Awrite(a.id, type, buf) < 0 ? -1 : 0;
}

// reads a string from an attribute or dataset element type `ftype`
int read_strings(hid_t obj, bool is_attr, hid_t ftype, size_t n, String1 &out)
{
    out.assign(n, std::string());
    if (n == 0) return 0;
    if (H5Tis_variable_str(ftype) > 0) {
        Tid mt;
        mt.id = vlen_string_type();
        std::vector<char *> buf(n, nullptr);
        const herr_t e = is_attr ? H5Aread(obj, mt.id, buf.data()) : H5Dread(obj, mt.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data());
        if (e < 0) return -1;
        for (size_t i = 0; i < n; i++) {
            if (buf[i]) {
                out[i] = buf[i];
                H5free_memory(buf[i]);
            }
        }
        return 0;
    }
    const size_t len = H5Tget_size(ftype);
    Tid mt;
    mt.id = H5Tcopy(H5T_C_S1);
    H5Tset_size(mt.id, len);
    std::vector<char> buf(n * len + 1, '\0');
    const herr_t e = is_attr ? H5Aread(obj, mt.id, buf.data()) : H5Dread(obj, mt.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data());
    if (e < 0) return -1;
    for (size_t i = 0; i < n; i++) out[i] = std::string(buf.data() + i * len, strnlen(buf.data() + i * len, len));
    return 0;
}

// dataset shape; false when absent
bool dataset_shape(hid_t file, const std::string &path, std::vector<hsize_t> &dims)
{
    if (!link_exists(file, path)) return false;
    H5O_info_t info;
    if (H5Oget_info_by_name(file, path.c_str(), &info, H5P_DEFAULT) < 0 || info.type != H5O_TYPE_DATASET)
        return false;
    Did d;
    if ((d.id = H5Dopen2(file, path.c_str(), H5P_DEFAULT)) < 0) return false;
    Sid s;
    s.id = H5Dget_space(d.id);
    const int rank = H5Sget_simple_extent_ndims(s.id);
    dims.assign(rank > 0 ? rank : 0, 0);
    if (rank > 0) H5Sget_simple_extent_dims(s.id, dims.data(), nullptr);
    return true;
}

herr_t count_matching(hid_t, const char *name, const H5L_info_t *, void *data)
{
    auto *c = static_cast<std::pair<std::string, size_t> *>(data);
    if (std::strstr(name, c->first.c_str())) c->second++;
    return 0;
}

// members of the group containing `name` whose name contains its last component
size_t count_members(hid_t file, const std::string &name)
{
    const size_t slash = name.find_last_of('/');
    const std::string group = slash == std::string::npos ? std::string("/") : strip(name.substr(0, slash + 1));
    const std::string object = slash == std::string::npos ? name : name.substr(slash + 1);
    if (group != "/" && !link_exists(file, group)) return 0;
    Gid g;
    if ((g.id = H5Oopen(file, group.c_str(), H5P_DEFAULT)) < 0) return 0;
    std::pair<std::string, size_t> c(object, 0);
    H5Literate(g.id, H5_INDEX_NAME, H5_ITER_INC, nullptr, count_matching, &c);
    return c.second;
}

template <typename T>
int read_data(const std::string &file_name, const std::string &name, std::vector<T> &flat, std::vector<hsize_t> &dims)
{
    Fid f;
    if ((f.id = open_read(file_name)) < 0) return -1;
    const std::string p = strip(name);
    if (!dataset_shape(f.id, p, dims)) return -1;
    size_t n = 1;
    for (hsize_t d : dims) n *= (size_t)d;
    flat.assign(n, T());
    Did d;
    if ((d.id = H5Dopen2(f.id, p.c_str(), H5P_DEFAULT)) < 0) return -1;
    if (n == 0) return 0;
    return H5Dread(d.id, native<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, flat.data()) < 0 ? -1 : 0;
}

// creates the dataset `name` (replacing an existing one) of shape dims and
// element type `type`, and writes `buf`
int write_data(const std::string &file_name, const std::string &name, hid_t type, const std::vector<hsize_t> &dims,
               const void *buf)
{
    Fid f;
    if ((f.id = open_write(file_name)) < 0) return -1;
    const std::string p = strip(name);
    if (link_exists(f.id, p) && H5Ldelete(f.id, p.c_str(), H5P_DEFAULT) < 0) return -1;
    Sid s;
    s.id = H5Screate_simple((int)dims.size(), dims.data(), nullptr);
    Pid lcpl;
    lcpl.id = link_create_plist();
    Did d;
    if ((d.id = H5Dcreate2(f.id, p.c_str(), type, s.id, lcpl.id, H5P_DEFAULT, H5P_DEFAULT)) < 0) return -1;
    return H5Dwrite(d.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0 ? -1 : 0;
}

template <typename T>
int read_list_t(const std::string &file_name, const std::string &name, std::vector<T> &list)
{
    std::vector<hsize_t> dims;
    return read_data(file_name, name, list, dims);
}

template <typename T>
int read_array_t(const std::string &file_name, const std::string &name, std::vector<std::vector<T>> &array)
{
    std::vector<T> flat;
    std::vector<hsize_t> dims;
    if (read_data(file_name, name, flat, dims)) return -1;
    if (dims.size() != 2) return -1;
    array.assign((size_t)dims[0], std::vector<T>((size_t)dims[1]));
    for (size_t i = 0; i < (size_t)dims[0]; i++)
        for (size_t j = 0; j < (size_t)dims[1]; j++) array[i][j] = flat[i * (size_t)dims[1] + j];
    return 0;
}

template <typename T>
int write_list_t(const std::string &file_name, const std::string &name, const std::vector<T> &list)
{
    if (list.empty()) return 0;
    return write_data(file_name, name, native<T>(), {(hsize_t)list.size()}, list.data());
}

template <typename T>
int write_array_t(const std::string &file_name, const std::string &name, const std::vector<std::vector<T>> &array)
{
    if (array.empty()) return 0;
    const size_t cols = array[0].size();
    std::vector<T> flat;
    flat.reserve(array.size() * cols);
    for (const auto &row : array) {
        if (row.size() != cols) return -1;   // HDF5 datasets are rectangular
        flat.insert(flat.end(), row.begin(), row.end());
    }
    return write_data(file_name, name, native<T>(), {(hsize_t)array.size(), (hsize_t)cols}, flat.data());
}

}  // namespace

int IoHDF5::read_length(const std::string &name, size_t &length) const
{
    length = 0;
    Fid f;
    if ((f.id = open_read(io_file)) < 0) return -1;
    std::vector<hsize_t> dims;
    if (dataset_shape(f.id, strip(name), dims)) {
        length = dims.empty() ? 1 : (size_t)dims[0];
        return 0;
    }
    length = count_members(f.id, name);
    return 0;
}

int IoHDF5::read_width(const std::string &name, size_t &width) const
{
    width = 0;
    Fid f;
    if ((f.id = open_read(io_file)) < 0) return -1;
    std::vector<hsize_t> dims;
    if (dataset_shape(f.id, strip(name), dims) && dims.size() >= 2) {
        width = (size_t)dims[1];
        return 0;
    }
    width = count_members(f.id, name);
    return 0;
}

int IoHDF5::read_number(const std::string &name, size_t &number) const { return read_attribute_number(io_file, name, number); }
int IoHDF5::read_number(const std::string &name, long &number) const { return read_attribute_number(io_file, name, number); }
int IoHDF5::read_number(const std::string &name, double &number) const { return read_attribute_number(io_file, name, number); }

int IoHDF5::read_word(const std::string &name, std::string &word) const
{
    std::string obj, attr;
    if (!split_attribute(name, obj, attr)) return -1;
    Fid f;
    if ((f.id = open_read(io_file)) < 0) return -1;
    if (obj != "/" && !link_exists(f.id, obj)) return -1;
    Gid g;
    if ((g.id = H5Oopen(f.id, obj.c_str(), H5P_DEFAULT)) < 0) return -1;
    if (H5Aexists(g.id, attr.c_str()) <= 0) return -1;
    Aid a;
    if ((a.id = H5Aopen(g.id, attr.c_str(), H5P_DEFAULT)) < 0) return -1;
    Tid t;
    t.id = H5Aget_type(a.id);
    if (H5Tget_class(t.id) != H5T_STRING) return -1;
    String1 one;
    if (read_strings(a.id, true, t.id, 1, one)) return -1;
    word = one[0];
    return 0;
}

// IoPython::read_bool (io_python.cpp:172-185): booleans are words
int IoHDF5::read_bool(const std::string &name, bool &value) const
{
    std::string word;
    int err = read_word(name, word);
    if (word == "true") value = true;
    else if (word == "false") value = false;
    else err = -1;
    return err;
}

IoPython::IoPython(const std::string &imp, const std::string &file)
    : IoHDF5(file), implementation("io_python_" + imp)
{
    if (imp != "hdf5")
        throw std::invalid_argument("IoPython: only the \"hdf5\" implementation is built (IoText covers text)");
}

int IoHDF5::read_list(const std::string &name, Long1 &list) const { return read_list_t(io_file, name, list); }
int IoHDF5::read_list(const std::string &name, Double1 &list) const { return read_list_t(io_file, name, list); }

int IoHDF5::read_list(const std::string &name, String1 &list) const
{
    Fid f;
    if ((f.id = open_read(io_file)) < 0) return -1;
    const std::string p = strip(name);
    std::vector<hsize_t> dims;
    if (!dataset_shape(f.id, p, dims)) return -1;
    size_t n = 1;
    for (hsize_t d : dims) n *= (size_t)d;
    Did d;
    if ((d.id = H5Dopen2(f.id, p.c_str(), H5P_DEFAULT)) < 0) return -1;
    Tid t;
    t.id = H5Dget_type(d.id);
    if (H5Tget_class(t.id) != H5T_STRING) return -1;
    return read_strings(d.id, false, t.id, n, list);
}

int IoHDF5::read_array(const std::string &name, Long2 &array) const { return read_array_t(io_file, name, array); }
int IoHDF5::read_array(const std::string &name, Double2 &array) const { return read_array_t(io_file, name, array); }

int IoHDF5::write_number(const std::string &name, const size_t &number) const
{
    const long v = (long)number;   // Python ints are stored as 64-bit signed integers
    return write_attribute(io_file, name, H5T_NATIVE_INT64, &v);
}
int IoHDF5::write_number(const std::string &name, const long &number) const
{
    return write_attribute(io_file, name, H5T_NATIVE_INT64, &number);
}
int IoHDF5::write_number(const std::string &name, const double &number) const
{
    return write_attribute(io_file, name, H5T_NATIVE_DOUBLE, &number);
}

int IoHDF5::write_word(const std::string &name, const std::string &word) const
{
    Tid t;
    t.id = vlen_string_type();
    const char *p = word.c_str();
    return write_attribute(io_file, name, t.id, &p);
}

int IoHDF5::write_bool(const std::string &name, const bool &value) const
{
    return write_word(name, value ? "true" : "false");
}

int IoHDF5::write_list(const std::string &name, const Long1 &list) const { return write_list_t(io_file, name, list); }
int IoHDF5::write_list(const std::string &name, const Double1 &list) const { return write_list_t(io_file, name, list); }

int IoHDF5::write_list(const std::string &name, const String1 &list) const
{
    if (list.empty()) return 0;
    Tid t;
    t.id = vlen_string_type();
    std::vector<const char *> ptr(list.size());
    for (size_t i = 0; i < list.size(); i++) ptr[i] = list[i].c_str();
    return write_data(io_file, name, t.id, {(hsize_t)list.size()}, ptr.data());
}

int IoHDF5::write_array(const std::string &name, const Long2 &array) const { return write_array_t(io_file, name, array); }
int IoHDF5::write_array(const std::string &name, const Double2 &array) const { return write_array_t(io_file, name, array); }

}  // namespace magritte
