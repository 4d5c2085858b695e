/*
 * H5MD file I/O for HAL's MD on MI355X: a host-only library
 * (libhalmd_h5md.so, linked against the HDF5 C library), kept apart from
 * libhalmd_hip.so so that the device library does not depend on HDF5.
 *
 * Mirrors the reference's writers and readers:
 *   halmd/io/writers/h5md/file.cpp:43-75, append.cpp:39-118, truncate.cpp:36-103
 *   halmd/io/readers/h5md/file.cpp:40-115, append.cpp:38-160, truncate.cpp
 * and the H5MD layout the Lua modules write (lua/halmd/observables/
 * phase_space.lua.in:236-275, thermodynamics.lua.in:213-284,
 * lua/halmd/mdsim/box.lua.in:138-170).
 *
 * Conventions: every function returns 0 on success or a HALMD_ERR_* code
 * (halmd_hip.h); halmd_h5md_last_error() gives the message.  Locations are
 * '/'-separated group paths below the file root ("particles/all").
 * Element types: HALMD_H5_F64, F32, I32, U32, I64, U64.
 */
#ifndef HALMD_H5MD_H
#define HALMD_H5MD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HALMD_H5_F64 0
#define HALMD_H5_F32 1
#define HALMD_H5_I32 2
#define HALMD_H5_U32 3
#define HALMD_H5_I64 4
#define HALMD_H5_U64 5

typedef struct halmd_h5md_file halmd_h5md_file;
typedef struct halmd_h5md_append halmd_h5md_append;

const char* halmd_h5md_last_error(void);

/* ---- file: writers/h5md/file.cpp:43-75 (file(path, author_name,
 * author_email, overwrite)): /h5md with attribute version {1, 0}, groups
 * creator (name, version) and author (name, email if given) ---- */
int halmd_h5md_file_create(const char* path, const char* author_name, const char* author_email, int overwrite,
                           halmd_h5md_file** out);
/* readers/h5md/file.cpp:40-115: open read-only and check the /h5md
 * attributes (version {1, 0}) */
int halmd_h5md_file_open(const char* path, halmd_h5md_file** out);
int halmd_h5md_file_flush(halmd_h5md_file* f);
int halmd_h5md_file_close(halmd_h5md_file* f);
/* /h5md version (2 ints) */
int halmd_h5md_file_version(halmd_h5md_file* f, int* version);
/* string attribute `name` of the object at `location` into out (cap bytes,
 * NUL-terminated) */
int halmd_h5md_read_attribute_string(halmd_h5md_file* f, const char* location, const char* name, char* out,
                                     size_t cap);
/* attributes on the group at `location` (created with its parents as
 * needed): integers, or n strings */
int halmd_h5md_write_attribute_int(halmd_h5md_file* f, const char* location, const char* name, const int* value,
                                   size_t n);
int halmd_h5md_write_attribute_strings(halmd_h5md_file* f, const char* location, const char* name,
                                       const char* const* value, size_t n);
int halmd_h5md_read_attribute_int(halmd_h5md_file* f, const char* location, const char* name, int* value,
                                  size_t n);
/* does `location` exist (1) or not (0) */
int halmd_h5md_exists(halmd_h5md_file* f, const char* location, int* out);

/* ---- truncate mode: writers/h5md/truncate.cpp:36-103 ----
 * dataset `name` in group `location`, (re)written whole: rank dims */
int halmd_h5md_truncate_write(halmd_h5md_file* f, const char* location, const char* name, int type, const void* data,
                              int rank, const uint64_t* dims);
/* readers/h5md/truncate.cpp: the whole dataset (shape must match) */
int halmd_h5md_truncate_read(halmd_h5md_file* f, const char* location, const char* name, int type, void* data,
                             int rank, const uint64_t* dims);
/* rank and dims (at most 8) of the dataset at `path` */
int halmd_h5md_dataset_shape(halmd_h5md_file* f, const char* path, int* rank, uint64_t* dims);

/* ---- append mode: writers/h5md/append.cpp:39-118 ----
 * A writer owns a group with one step and one time dataset; each field is a
 * subgroup (location/name) holding a chunked `value` dataset with a first
 * dimension that grows by one sample per write, and hard links `step` and
 * `time` to the writer's shared datasets. */
int halmd_h5md_append_create(halmd_h5md_file* f, const char* location, halmd_h5md_append** out);
int halmd_h5md_append_destroy(halmd_h5md_append* w);
/* start a sample: append step and time (append::write_step_time) */
int halmd_h5md_append_begin(halmd_h5md_append* w, int64_t step, double time);
/* append one sample of field `name` (a path below the writer's group) with
 * per-sample shape (rank, dims); every field writes once per sample, and a
 * field's first write must come with the writer's first sample */
int halmd_h5md_append_field(halmd_h5md_append* w, const char* name, int type, const void* data, int rank,
                            const uint64_t* dims);

/* ---- append reader: readers/h5md/append.cpp:38-160 ---- */
/* samples of the field group at `location` */
int halmd_h5md_append_samples(halmd_h5md_file* f, const char* location, uint64_t* count);
/* index of the sample at step `offset` (a negative offset counts from the
 * last step + 1: -1 is the last; read_step_index) or at time `offset` (a
 * negative offset counts from the last time; read_time_index, with the
 * reference's relative tolerance of 100 epsilon) */
int halmd_h5md_append_index_at_step(halmd_h5md_file* f, const char* location, int64_t offset, uint64_t* index);
int halmd_h5md_append_index_at_time(halmd_h5md_file* f, const char* location, double offset, uint64_t* index);
/* step and time of sample `index` */
int halmd_h5md_append_step_time(halmd_h5md_file* f, const char* location, uint64_t index, int64_t* step,
                                double* time);
/* sample `index` of the field's value (per-sample shape must match) */
int halmd_h5md_append_read(halmd_h5md_file* f, const char* location, uint64_t index, int type, void* data, int rank,
                           const uint64_t* dims);

#ifdef __cplusplus
}
#endif

#endif
