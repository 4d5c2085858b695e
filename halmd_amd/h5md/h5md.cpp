// H5MD file I/O over the HDF5 C library (host only).
//
// The layout follows the reference's writers (halmd/io/writers/h5md/
// {file,append,truncate}.cpp) and readers (halmd/io/readers/h5md/
// {file,append,truncate}.cpp):
//  - /h5md: attribute version {1, 0}; groups creator {name, version} and
//    author {name, email};
//  - append mode: a writer group holds one step (int64) and one time
//    (double) dataset, both without a name of their own (created anonymous,
//    as the reference unlinks them after creation); every field is a
//    subgroup with a chunked `value` dataset whose first dimension grows by
//    one sample per write, one sample per chunk, and hard links `step` and
//    `time` to the shared datasets;
//  - truncate mode: a dataset written whole, rewritten in place.
#include "../../include/halmd_h5md.h"

#include <hdf5.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr int ERR_INVALID_ARGUMENT = 1;   // HALMD_ERR_* of halmd_hip.h
constexpr int ERR_LOGIC = 4;
constexpr int ERR_IO = 5;

struct error : std::runtime_error {
    int code;
    error(int c, std::string const& m) : std::runtime_error(m), code(c) {}
};

thread_local std::string last_error;

template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (error const& e) {
        last_error = e.what();
        return e.code;
    } catch (std::exception const& e) {
        last_error = e.what();
        return ERR_LOGIC;
    }
}

void require(bool c, char const* msg) {
    if (!c) throw error(ERR_INVALID_ARGUMENT, msg);
}

template <class T>
T check(T r, std::string const& what) {
    if (r < 0) throw error(ERR_IO, "HDF5: " + what);
    return r;
}

// owned HDF5 identifier, closed with the matching H5*close
struct hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    hid() = default;
    hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    hid(hid const&) = delete;
    hid& operator=(hid const&) = delete;
    hid(hid&& o) noexcept : id(o.id), close(o.close) { o.id = -1; }
    hid& operator=(hid&& o) noexcept {
        reset();
        id = o.id;
        close = o.close;
        o.id = -1;
        return *this;
    }
    ~hid() { reset(); }
    void reset() {
        if (id >= 0 && close) close(id);
        id = -1;
    }
    operator hid_t() const { return id; }
};

hid group_id(hid_t g) { return hid(g, H5Gclose); }
hid dataset_id(hid_t d) { return hid(d, H5Dclose); }
hid space_id(hid_t s) { return hid(s, H5Sclose); }
hid type_id(hid_t t) { return hid(t, H5Tclose); }
hid plist_id(hid_t p) { return hid(p, H5Pclose); }
hid attr_id(hid_t a) { return hid(a, H5Aclose); }

void quiet_hdf5() {
    static bool done = false;
    if (!done) {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        done = true;
    }
}

std::vector<std::string> split(std::string const& path) {
    std::vector<std::string> out;
    size_t b = 0;
    while (b <= path.size()) {
        size_t e = path.find('/', b);
        if (e == std::string::npos) e = path.size();
        if (e > b) out.push_back(path.substr(b, e - b));
        b = e + 1;
    }
    return out;
}

bool link_exists(hid_t loc, std::string const& name) {
    htri_t const r = H5Lexists(loc, name.c_str(), H5P_DEFAULT);
    return r > 0;
}

// the group at `path` below loc (h5xx::open_group: created with its parents
// when create is set)
hid open_group(hid_t loc, std::string const& path, bool create) {
    hid cur = group_id(H5Iget_type(loc) == H5I_GROUP ? check(H5Oopen(loc, ".", H5P_DEFAULT), "open group")
                                                      : check(H5Gopen2(loc, "/", H5P_DEFAULT), "open root group"));
    for (auto const& c : split(path)) {
        if (link_exists(cur, c)) {
            cur = group_id(check(H5Gopen2(cur, c.c_str(), H5P_DEFAULT), "open group " + c));
        } else {
            if (!create) throw error(ERR_INVALID_ARGUMENT, "no group '" + c + "' in location '" + path + "'");
            cur = group_id(check(H5Gcreate2(cur, c.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                                 "create group " + c));
        }
    }
    return cur;
}

hid_t mem_type(int type) {
    switch (type) {
    case HALMD_H5_F64: return H5T_NATIVE_DOUBLE;
    case HALMD_H5_F32: return H5T_NATIVE_FLOAT;
    case HALMD_H5_I32: return H5T_NATIVE_INT32;
    case HALMD_H5_U32: return H5T_NATIVE_UINT32;
    case HALMD_H5_I64: return H5T_NATIVE_INT64;
    case HALMD_H5_U64: return H5T_NATIVE_UINT64;
    }
    throw error(ERR_INVALID_ARGUMENT, "h5md: unknown element type");
}

std::vector<hsize_t> dims_of(int rank, uint64_t const* dims) {
    require(rank >= 0 && rank <= 7, "h5md: rank must be in 0 .. 7");
    require(rank == 0 || dims != nullptr, "h5md: dims required");
    return std::vector<hsize_t>(dims, dims + rank);
}

std::vector<hsize_t> extent(hid_t dataset) {
    hid s = space_id(check(H5Dget_space(dataset), "dataset space"));
    int const r = H5Sget_simple_extent_ndims(s);
    check(r, "dataset rank");
    std::vector<hsize_t> d(static_cast<size_t>(r));
    if (r) check(H5Sget_simple_extent_dims(s, d.data(), nullptr), "dataset dims");
    return d;
}

hid create_chunked(hid_t loc, char const* name, hid_t file_type, std::vector<hsize_t> const& sample) {
    std::vector<hsize_t> d{0}, m{H5S_UNLIMITED}, c{1};
    for (hsize_t x : sample) {
        d.push_back(x);
        m.push_back(x);
        c.push_back(std::max<hsize_t>(x, 1));
    }
    hid space = space_id(check(H5Screate_simple(int(d.size()), d.data(), m.data()), "dataspace"));
    hid dcpl = plist_id(check(H5Pcreate(H5P_DATASET_CREATE), "plist"));
    check(H5Pset_chunk(dcpl, int(c.size()), c.data()), "chunk");
    if (name)
        return dataset_id(check(H5Dcreate2(loc, name, file_type, space, H5P_DEFAULT, dcpl, H5P_DEFAULT),
                                std::string("create dataset ") + name));
    return dataset_id(check(H5Dcreate_anon(loc, file_type, space, dcpl, H5P_DEFAULT), "create dataset"));
}

// append one sample (h5xx::write_chunked_dataset)
void append_sample(hid_t dataset, hid_t mtype, void const* data) {
    std::vector<hsize_t> d = extent(dataset);
    require(!d.empty(), "h5md: not an append dataset");
    std::vector<hsize_t> start(d.size(), 0), count = d;
    start[0] = d[0];
    count[0] = 1;
    d[0] += 1;
    check(H5Dset_extent(dataset, d.data()), "extend dataset");
    hid fs = space_id(check(H5Dget_space(dataset), "dataset space"));
    check(H5Sselect_hyperslab(fs, H5S_SELECT_SET, start.data(), nullptr, count.data(), nullptr), "select");
    hid ms = space_id(check(H5Screate_simple(int(count.size()), count.data(), nullptr), "dataspace"));
    check(H5Dwrite(dataset, mtype, ms, fs, H5P_DEFAULT, data), "write sample");
}

// read sample `index` (h5xx::read_chunked_dataset)
void read_sample(hid_t dataset, hid_t mtype, uint64_t index, void* data, std::vector<hsize_t> const& sample) {
    std::vector<hsize_t> d = extent(dataset);
    require(d.size() == sample.size() + 1, "h5md: sample rank does not match the dataset");
    require(std::equal(sample.begin(), sample.end(), d.begin() + 1), "h5md: sample shape does not match the dataset");
    require(index < d[0], "h5md: sample index out of range");
    std::vector<hsize_t> start(d.size(), 0), count = d;
    start[0] = index;
    count[0] = 1;
    hid fs = space_id(check(H5Dget_space(dataset), "dataset space"));
    check(H5Sselect_hyperslab(fs, H5S_SELECT_SET, start.data(), nullptr, count.data(), nullptr), "select");
    hid ms = space_id(check(H5Screate_simple(int(count.size()), count.data(), nullptr), "dataspace"));
    check(H5Dread(dataset, mtype, ms, fs, H5P_DEFAULT, data), "read sample");
}

template <class T>
std::vector<T> read_all_1d(hid_t dataset, hid_t mtype) {
    std::vector<hsize_t> d = extent(dataset);
    require(d.size() == 1, "h5md: expected a one-dimensional dataset");
    std::vector<T> v(d[0]);
    if (!v.empty()) check(H5Dread(dataset, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()), "read dataset");
    return v;
}

hid fixed_string_type(size_t len) {
    hid t = type_id(check(H5Tcopy(H5T_C_S1), "string type"));
    check(H5Tset_size(t, std::max<size_t>(len, 1) + 1), "string size");
    check(H5Tset_strpad(t, H5T_STR_NULLTERM), "string padding");
    return t;
}

void delete_attribute_if(hid_t obj, char const* name) {
    if (H5Aexists(obj, name) > 0) check(H5Adelete(obj, name), "delete attribute");
}

void write_string_attribute(hid_t obj, char const* name, std::string const& value) {
    delete_attribute_if(obj, name);
    hid t = fixed_string_type(value.size());
    hid s = space_id(check(H5Screate(H5S_SCALAR), "dataspace"));
    hid a = attr_id(check(H5Acreate2(obj, name, t, s, H5P_DEFAULT, H5P_DEFAULT), std::string("attribute ") + name));
    std::vector<char> buf(value.size() + 1, 0);
    std::memcpy(buf.data(), value.data(), value.size());
    check(H5Awrite(a, t, buf.data()), "write attribute");
}

} // namespace

struct halmd_h5md_file {
    hid_t id = -1;
    bool writable = false;
    std::string path;
};

struct halmd_h5md_append {
    halmd_h5md_file* f = nullptr;
    hid group, step, time;
    uint64_t samples = 0;
    struct field {
        hid group, value;
        int type;
        std::vector<hsize_t> sample;
        uint64_t written = 0;
    };
    std::map<std::string, field> fields;
};

extern "C" {

const char* halmd_h5md_last_error(void) { return last_error.c_str(); }

int halmd_h5md_file_create(const char* path, const char* author_name, const char* author_email, int overwrite,
                           halmd_h5md_file** out) {
    return guard([&] {
        quiet_hdf5();
        require(path && out, "h5md: path and out required");
        if (access(path, F_OK) == 0 && !overwrite)
            throw error(ERR_INVALID_ARGUMENT, std::string("Output file \"") + path + "\" exists and won't be overwritten");
        hid_t const id = check(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), std::string("create ") + path);
        auto f = std::make_unique<halmd_h5md_file>();
        f->id = id;
        f->writable = true;
        f->path = path;
        hid h5md = open_group(id, "h5md", true);
        int const version[2] = {1, 0};
        hsize_t two = 2;
        {
            hid s = space_id(check(H5Screate_simple(1, &two, nullptr), "dataspace"));
            hid a = attr_id(check(H5Acreate2(h5md, "version", H5T_NATIVE_INT32, s, H5P_DEFAULT, H5P_DEFAULT),
                                  "attribute version"));
            check(H5Awrite(a, H5T_NATIVE_INT32, version), "write version");
        }
        hid creator = open_group(h5md, "creator", true);
        write_string_attribute(creator, "name", "HAL's MD package (MI355X)");
        write_string_attribute(creator, "version", "1.0");
        hid author = open_group(h5md, "author", true);
        std::string name = author_name && *author_name ? author_name : "";
        if (name.empty()) {
            char const* u = std::getenv("USER");
            name = u ? u : "unknown";
        }
        write_string_attribute(author, "name", name);
        if (author_email && *author_email) write_string_attribute(author, "email", author_email);
        *out = f.release();
    });
}

int halmd_h5md_file_open(const char* path, halmd_h5md_file** out) {
    return guard([&] {
        quiet_hdf5();
        require(path && out, "h5md: path and out required");
        hid_t const id = check(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), std::string("open ") + path);
        auto f = std::make_unique<halmd_h5md_file>();
        f->id = id;
        f->path = path;
        // readers/h5md/file.cpp:58-115: the /h5md group and its attributes
        if (!link_exists(id, "h5md")) {
            H5Fclose(id);
            f->id = -1;
            throw error(ERR_INVALID_ARGUMENT, "not an H5MD file: no /h5md group");
        }
        int v[2] = {0, 0};
        hid h5md = group_id(check(H5Gopen2(id, "h5md", H5P_DEFAULT), "open /h5md"));
        if (H5Aexists(h5md, "version") <= 0) throw error(ERR_INVALID_ARGUMENT, "H5MD: no version attribute");
        hid a = attr_id(check(H5Aopen(h5md, "version", H5P_DEFAULT), "open version"));
        check(H5Aread(a, H5T_NATIVE_INT32, v), "read version");
        if (v[0] != 1 || v[1] != 0)
            throw error(ERR_INVALID_ARGUMENT,
                        "version of H5MD format: " + std::to_string(v[0]) + "." + std::to_string(v[1]));
        *out = f.release();
    });
}

int halmd_h5md_file_flush(halmd_h5md_file* f) {
    return guard([&] {
        require(f, "h5md: null file");
        check(H5Fflush(f->id, H5F_SCOPE_GLOBAL), "flush");
    });
}

int halmd_h5md_file_close(halmd_h5md_file* f) {
    return guard([&] {
        if (!f) return;
        std::unique_ptr<halmd_h5md_file> own(f);
        if (f->id >= 0) check(H5Fclose(f->id), "close");
    });
}

int halmd_h5md_file_version(halmd_h5md_file* f, int* version) {
    return halmd_h5md_read_attribute_int(f, "h5md", "version", version, 2);
}

int halmd_h5md_exists(halmd_h5md_file* f, const char* location, int* out) {
    return guard([&] {
        require(f && location && out, "h5md: null argument");
        hid_t cur = f->id;
        std::vector<hid> keep;
        *out = 0;
        auto parts = split(location);
        for (size_t k = 0; k < parts.size(); ++k) {
            if (!link_exists(cur, parts[k])) return;
            if (k + 1 < parts.size()) {
                keep.push_back(group_id(check(H5Gopen2(cur, parts[k].c_str(), H5P_DEFAULT), "open group")));
                cur = keep.back();
            }
        }
        *out = 1;
    });
}

int halmd_h5md_read_attribute_string(halmd_h5md_file* f, const char* location, const char* name, char* out,
                                     size_t cap) {
    return guard([&] {
        require(f && location && name && out && cap > 0, "h5md: null argument");
        hid obj = hid(check(H5Oopen(f->id, *location ? location : "/", H5P_DEFAULT), std::string("open ") + location),
                      H5Oclose);
        if (H5Aexists(obj, name) <= 0) throw error(ERR_INVALID_ARGUMENT, std::string("no attribute ") + name);
        hid a = attr_id(check(H5Aopen(obj, name, H5P_DEFAULT), "open attribute"));
        hid t = type_id(check(H5Aget_type(a), "attribute type"));
        require(H5Tget_class(t) == H5T_STRING, "h5md: attribute is not a string");
        std::string value;
        if (H5Tis_variable_str(t) > 0) {
            char* p = nullptr;
            hid mt = type_id(check(H5Tcopy(H5T_C_S1), "string type"));
            check(H5Tset_size(mt, H5T_VARIABLE), "string size");
            check(H5Aread(a, mt, &p), "read attribute");
            value = p ? p : "";
            hid s = space_id(check(H5Aget_space(a), "attribute space"));
            H5Dvlen_reclaim(mt, s, H5P_DEFAULT, &p);
        } else {
            size_t const n = H5Tget_size(t);
            std::vector<char> buf(n + 1, 0);
            check(H5Aread(a, t, buf.data()), "read attribute");
            value = buf.data();
        }
        require(value.size() < cap, "h5md: attribute longer than the buffer");
        std::memcpy(out, value.c_str(), value.size() + 1);
    });
}

int halmd_h5md_write_attribute_int(halmd_h5md_file* f, const char* location, const char* name, const int* value,
                                   size_t n) {
    return guard([&] {
        require(f && f->writable && location && name && value && n > 0, "h5md: invalid argument");
        hid g = open_group(f->id, location, true);
        delete_attribute_if(g, name);
        hsize_t d = n;
        hid s = space_id(n == 1 ? check(H5Screate(H5S_SCALAR), "dataspace")
                                : check(H5Screate_simple(1, &d, nullptr), "dataspace"));
        hid a = attr_id(check(H5Acreate2(g, name, H5T_NATIVE_INT32, s, H5P_DEFAULT, H5P_DEFAULT), "attribute"));
        check(H5Awrite(a, H5T_NATIVE_INT32, value), "write attribute");
    });
}

int halmd_h5md_read_attribute_int(halmd_h5md_file* f, const char* location, const char* name, int* value,
                                  size_t n) {
    return guard([&] {
        require(f && location && name && value && n > 0, "h5md: invalid argument");
        hid obj = hid(check(H5Oopen(f->id, *location ? location : "/", H5P_DEFAULT), std::string("open ") + location),
                      H5Oclose);
        if (H5Aexists(obj, name) <= 0) throw error(ERR_INVALID_ARGUMENT, std::string("no attribute ") + name);
        hid a = attr_id(check(H5Aopen(obj, name, H5P_DEFAULT), "open attribute"));
        hid s = space_id(check(H5Aget_space(a), "attribute space"));
        hssize_t const m = H5Sget_simple_extent_npoints(s);
        require(m == hssize_t(n), "h5md: attribute size mismatch");
        check(H5Aread(a, H5T_NATIVE_INT32, value), "read attribute");
    });
}

int halmd_h5md_write_attribute_strings(halmd_h5md_file* f, const char* location, const char* name,
                                       const char* const* value, size_t n) {
    return guard([&] {
        require(f && f->writable && location && name && value && n > 0, "h5md: invalid argument");
        hid g = open_group(f->id, location, true);
        delete_attribute_if(g, name);
        size_t len = 0;
        for (size_t k = 0; k < n; ++k) len = std::max(len, std::strlen(value[k]));
        hid t = fixed_string_type(len);
        std::vector<char> buf(n * (len + 1), 0);
        for (size_t k = 0; k < n; ++k) std::memcpy(&buf[k * (len + 1)], value[k], std::strlen(value[k]));
        hsize_t d = n;
        hid s = space_id(check(H5Screate_simple(1, &d, nullptr), "dataspace"));
        hid a = attr_id(check(H5Acreate2(g, name, t, s, H5P_DEFAULT, H5P_DEFAULT), "attribute"));
        check(H5Awrite(a, t, buf.data()), "write attribute");
    });
}

int halmd_h5md_truncate_write(halmd_h5md_file* f, const char* location, const char* name, int type, const void* data,
                              int rank, const uint64_t* dims) {
    return guard([&] {
        require(f && f->writable && location && name && data, "h5md: invalid argument");
        hid_t const mt = mem_type(type);
        std::vector<hsize_t> d = dims_of(rank, dims);
        hid g = open_group(f->id, location, true);
        hid ds;
        if (link_exists(g, name)) {
            ds = dataset_id(check(H5Dopen2(g, name, H5P_DEFAULT), std::string("open dataset ") + name));
            require(extent(ds) == d, "h5md: truncate write with another shape than the dataset's");
        } else {
            hid s = space_id(rank == 0 ? check(H5Screate(H5S_SCALAR), "dataspace")
                                       : check(H5Screate_simple(rank, d.data(), nullptr), "dataspace"));
            ds = dataset_id(check(H5Dcreate2(g, name, mt, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                                  std::string("create dataset ") + name));
        }
        check(H5Dwrite(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "write dataset");
    });
}

int halmd_h5md_truncate_read(halmd_h5md_file* f, const char* location, const char* name, int type, void* data,
                             int rank, const uint64_t* dims) {
    return guard([&] {
        require(f && location && name && data, "h5md: invalid argument");
        hid_t const mt = mem_type(type);
        std::vector<hsize_t> d = dims_of(rank, dims);
        hid g = open_group(f->id, location, false);
        hid ds = dataset_id(check(H5Dopen2(g, name, H5P_DEFAULT), std::string("open dataset ") + name));
        require(extent(ds) == d, "h5md: dataset shape does not match");
        check(H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "read dataset");
    });
}

int halmd_h5md_dataset_shape(halmd_h5md_file* f, const char* path, int* rank, uint64_t* dims) {
    return guard([&] {
        require(f && path && rank && dims, "h5md: invalid argument");
        hid ds = dataset_id(check(H5Dopen2(f->id, path, H5P_DEFAULT), std::string("open dataset ") + path));
        std::vector<hsize_t> d = extent(ds);
        require(d.size() <= 8, "h5md: rank above 8");
        *rank = int(d.size());
        for (size_t k = 0; k < d.size(); ++k) dims[k] = d[k];
    });
}

int halmd_h5md_append_create(halmd_h5md_file* f, const char* location, halmd_h5md_append** out) {
    return guard([&] {
        require(f && f->writable && location && *location && out, "h5md: invalid argument");
        auto w = std::make_unique<halmd_h5md_append>();
        w->f = f;
        w->group = open_group(f->id, location, true);
        w->step = create_chunked(w->group, nullptr, H5T_NATIVE_INT64, {});
        w->time = create_chunked(w->group, nullptr, H5T_NATIVE_DOUBLE, {});
        *out = w.release();
    });
}

int halmd_h5md_append_destroy(halmd_h5md_append* w) {
    return guard([&] { delete w; });
}

int halmd_h5md_append_begin(halmd_h5md_append* w, int64_t step, double time) {
    return guard([&] {
        require(w, "h5md: null writer");
        for (auto const& kv : w->fields)
            require(kv.second.written == w->samples, "h5md: a field was not written with the previous sample");
        append_sample(w->step, H5T_NATIVE_INT64, &step);
        append_sample(w->time, H5T_NATIVE_DOUBLE, &time);
        ++w->samples;
    });
}

int halmd_h5md_append_field(halmd_h5md_append* w, const char* name, int type, const void* data, int rank,
                            const uint64_t* dims) {
    return guard([&] {
        require(w && name && *name && data, "h5md: invalid argument");
        require(w->samples > 0, "h5md: append_begin first");
        hid_t const mt = mem_type(type);
        std::vector<hsize_t> d = dims_of(rank, dims);
        auto it = w->fields.find(name);
        if (it == w->fields.end()) {
            // append::on_write: the field group, its value dataset and links to
            // the shared step and time
            require(w->samples == 1, "h5md: a field must be written from the writer's first sample on");
            halmd_h5md_append::field fl;
            fl.group = open_group(w->group, name, true);
            fl.value = create_chunked(fl.group, "value", mt, d);
            check(H5Olink(w->step, fl.group, "step", H5P_DEFAULT, H5P_DEFAULT), "link step");
            check(H5Olink(w->time, fl.group, "time", H5P_DEFAULT, H5P_DEFAULT), "link time");
            fl.type = type;
            fl.sample = d;
            it = w->fields.emplace(name, std::move(fl)).first;
        }
        auto& fl = it->second;
        require(fl.type == type && fl.sample == d, "h5md: field written with another type or shape");
        require(fl.written + 1 == w->samples, "h5md: field written twice in one sample");
        append_sample(fl.value, mt, data);
        ++fl.written;
    });
}

int halmd_h5md_append_samples(halmd_h5md_file* f, const char* location, uint64_t* count) {
    return guard([&] {
        require(f && location && count, "h5md: invalid argument");
        hid g = open_group(f->id, location, false);
        hid ds = dataset_id(check(H5Dopen2(g, "step", H5P_DEFAULT), "open step"));
        *count = extent(ds).at(0);
    });
}

int halmd_h5md_append_index_at_step(halmd_h5md_file* f, const char* location, int64_t offset, uint64_t* index) {
    return guard([&] {
        require(f && location && index, "h5md: invalid argument");
        hid g = open_group(f->id, location, false);
        hid ds = dataset_id(check(H5Dopen2(g, "step", H5P_DEFAULT), "open step"));
        std::vector<int64_t> steps = read_all_1d<int64_t>(ds, H5T_NATIVE_INT64);
        if (steps.empty()) throw error(ERR_INVALID_ARGUMENT, "empty step dataset");
        int64_t const step = offset < 0 ? offset + steps.back() + 1 : offset;
        auto r = std::equal_range(steps.begin(), steps.end(), step);
        if (r.first == r.second) throw error(ERR_INVALID_ARGUMENT, "nonexistent step " + std::to_string(step));
        if (r.second - r.first > 1) throw error(ERR_INVALID_ARGUMENT, "ambiguous step " + std::to_string(step));
        *index = uint64_t(r.first - steps.begin());
    });
}

int halmd_h5md_append_index_at_time(halmd_h5md_file* f, const char* location, double offset, uint64_t* index) {
    return guard([&] {
        require(f && location && index, "h5md: invalid argument");
        hid g = open_group(f->id, location, false);
        hid ds = dataset_id(check(H5Dopen2(g, "time", H5P_DEFAULT), "open time"));
        std::vector<double> times = read_all_1d<double>(ds, H5T_NATIVE_DOUBLE);
        if (times.empty()) throw error(ERR_INVALID_ARGUMENT, "empty time dataset");
        double const time = std::signbit(offset) ? offset + times.back() : offset;
        auto less = [](double a, double b) {
            return a * (1 + 100 * std::numeric_limits<double>::epsilon()) < b;
        };
        auto r = std::equal_range(times.begin(), times.end(), time, less);
        if (r.first == r.second) throw error(ERR_INVALID_ARGUMENT, "nonexistent time");
        if (r.second - r.first > 1) throw error(ERR_INVALID_ARGUMENT, "ambiguous time");
        *index = uint64_t(r.first - times.begin());
    });
}

int halmd_h5md_append_step_time(halmd_h5md_file* f, const char* location, uint64_t index, int64_t* step,
                                double* time) {
    return guard([&] {
        require(f && location, "h5md: invalid argument");
        hid g = open_group(f->id, location, false);
        if (step) {
            hid ds = dataset_id(check(H5Dopen2(g, "step", H5P_DEFAULT), "open step"));
            read_sample(ds, H5T_NATIVE_INT64, index, step, {});
        }
        if (time) {
            hid ds = dataset_id(check(H5Dopen2(g, "time", H5P_DEFAULT), "open time"));
            read_sample(ds, H5T_NATIVE_DOUBLE, index, time, {});
        }
    });
}

int halmd_h5md_append_read(halmd_h5md_file* f, const char* location, uint64_t index, int type, void* data, int rank,
                           const uint64_t* dims) {
    return guard([&] {
        require(f && location && data, "h5md: invalid argument");
        hid_t const mt = mem_type(type);
        hid g = open_group(f->id, location, false);
        hid ds = dataset_id(check(H5Dopen2(g, "value", H5P_DEFAULT), "open value"));
        read_sample(ds, mt, index, data, dims_of(rank, dims));
    });
}

} // extern "C"
