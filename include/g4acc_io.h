/* g4acc_io: G4 files in the reference's HDF5 format, C ABI (host code; libg4acc_io.so).
 *
 * The reference writes a func::function<std::complex<double>, ...> with HDF5Writer::execute
 * (include/dca/io/hdf5/hdf5_writer.hpp:244-262, src/io/hdf5/hdf5_writer.cpp:153-193), which
 * DcaData uses for G4 in the group "functions" under the name "G4_<channel>"
 * (include/dca/phys/dca_data/dca_data.hpp:331-336):
 *   - element type: compound { "r": IEEE F64, "i": IEEE F64 } (hdf5_types.hpp:180-192);
 *   - dataspace: the function's dimensions reversed, slowest first (HDF5 is row-major, the
 *     function column-major, hdf5_writer.hpp:254-255);
 *   - attribute "domain-sizes": uint64 array of the dimensions, first index fastest;
 *   - attribute "name": a fixed-length C string of the function's name.
 * Values are interleaved (re, im) doubles in the function's order (first index fastest), which is
 * the order g4acc_get_g4 returns. Every call returns 0 on success and a negative code on error;
 * g4io_last_error() says why. No call throws.
 */
#ifndef G4ACC_IO_H
#define G4ACC_IO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* HDF5Writer::execute(name, f) for a complex function (hdf5_writer.hpp:244-262): writes the
 * dataset <group>/<name> of file `path` (created if missing, or truncated when truncate != 0),
 * with dims[0..ndims) the function's dimensions, first index fastest. An existing dataset of the
 * same shape is overwritten (hdf5_writer.cpp:155-167); a different shape is an error. */
int g4io_write_function(const char* path, const char* group, const char* name, const double* values,
                        const unsigned long long* dims, int ndims, int truncate);

/* HDF5Reader::execute for a complex function: reads dataset `dataset_path` (for example
 * "/functions/G4_PARTICLE_HOLE_CHARGE") into values (n_complex interleaved complex doubles,
 * first index fastest). dims (capacity max_dims) receives the dimensions first index fastest and
 * *ndims their count. values may be NULL to query the dimensions only. */
int g4io_read_function(const char* path, const char* dataset_path, double* values, size_t n_complex,
                       unsigned long long* dims, int max_dims, int* ndims);

/* The "name" attribute of a dataset (what HDF5Writer stores as f.get_name()). */
int g4io_read_name(const char* path, const char* dataset_path, char* out, size_t capacity);

/* FourPointType name of a channel value (four_point_type.cpp:37-56), e.g. 3 ->
 * "PARTICLE_HOLE_CHARGE"; NULL for an unknown value. */
const char* g4io_channel_name(int channel);

/* Message of the last failed call on this thread. */
const char* g4io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* G4ACC_IO_H */
