/*
 * odhost.h — C ABI of the host layer (libodhost.so): config.toml and the
 * solver-native HDF5 state files.  Plain C, status 0 = ok, message from
 * odh_last_error() otherwise.
 *
 * What each entry point replaces in drunkseal/opendynamo (file:line):
 *   odh_save_state      io.rs:59-136   save_state_h5 (+ parse_compression io.rs:35-52)
 *   odh_load_state      io.rs:163-234  load_state_h5 (validating loader), reading
 *                                      the native type (f64 resume is bit-exact)
 *   odh_read_header     io.rs:173-185  the attribute half of the loader
 *   odh_config_json     config.rs:205-209 load_config, returned as JSON text
 */
#ifndef ODHOST_H
#define ODHOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Write <out_dir>/<prefix>_<kind>_<label:06>.h5; kind "f32"/"f64";
 * compression as the [io] compression key.  path_out receives the path. */
int odh_save_state(uint64_t label, double t, const char* kind, const char* out_dir, const char* prefix,
                   const void* data, uint64_t nx, uint64_t ny, uint64_t nz, double lx, double ly, double lz,
                   const char* compression, char* path_out, size_t path_cap);

/* Validate mesh / fr_order / u_real like the reference loader and copy the 5
 * planes (native type) into data_out (bytes must match). */
int odh_load_state(const char* path, uint64_t nx, uint64_t ny, uint64_t nz, const char* kind, void* data_out,
                   size_t bytes, uint64_t* step, double* t);

/* Header without validation: u64[5] = nx ny nz fr_order step, f64[4] = t lx ly lz,
 * kind_out (>= 4 bytes) = u_real. */
int odh_read_header(const char* path, uint64_t u64_out[5], double f64_out[4], char* kind_out);

/* Parse a config.toml (reference keys and defaults) into JSON text. */
int odh_config_json(const char* path, char* json_out, size_t cap);

const char* odh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
