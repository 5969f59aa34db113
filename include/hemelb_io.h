/* Extraction (.xtr v5) and offset (.off v1) files: property output and the
 * binary64 distribution checkpoint of HemeLB.
 *
 * Replaces, on the host side of one rank:
 *   extraction::LocalPropertyOutput        (Code/extraction/LocalPropertyOutput.cc:52-360)
 *   extraction::LocalDistributionInput     (Code/extraction/LocalDistributionInput.cc:38-200)
 *   io::formats::extraction / offset       (Code/io/formats/extraction.h:10-80,
 *                                           Code/io/formats/offset.h:15-60)
 * All numbers are XDR (big-endian).  A file holds a 60-byte main header, the
 * field headers, then one record per written time step: the uint64 time
 * step followed, per site, by its uint32 x, y, z and every field's values.
 */
#ifndef HEMELB_IO_H
#define HEMELB_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field value types (extraction.h:37-45, TypeCode). */
enum hlbx_type { HLBX_FLOAT = 0, HLBX_DOUBLE = 1, HLBX_INT32 = 2, HLBX_UINT32 = 3,
                 HLBX_INT64 = 4, HLBX_UINT64 = 5 };

typedef struct {
  const char* name;
  int32_t type;           /* hlbx_type */
  int32_t n_elements;     /* values per site (1 pressure, 3 velocity, Q
                             distributions ...) */
  int32_t n_offsets;      /* 0, 1 or n_elements */
  const double* offsets;  /* stored in the field header and subtracted from
                             the values before they are written
                             (LocalPropertyOutput.cc:282-284) */
} hlbx_field;

typedef struct hlbx_writer hlbx_writer;

/* Creates `path` and its offset file (the same name with the extension
 * replaced by ".off", offset.h:51-56) for one rank holding `n_sites` sites
 * at `coords` [n][3]; fails if either file exists (the reference opens both
 * with MPI_MODE_EXCL).  Writes both headers. */
int hlbx_create(const char* path, double voxel_size_m, const double origin_m[3],
                int64_t n_sites, const int32_t* coords, int32_t n_fields,
                const hlbx_field* fields, hlbx_writer** out);
/* Appends the record of one time step; data[k] is double [n_sites][n_elements
 * of field k], converted to the field's type. */
int hlbx_append(hlbx_writer* w, uint64_t time_step, const double* const* data);
int hlbx_close(hlbx_writer* w);

/* Reads a distribution checkpoint (one DOUBLE field of q elements,
 * LocalDistributionInput.cc:38-170): the record of `target_time`, or the
 * last record when target_time < 0.  Sites are matched by coordinates and
 * returned in the caller's order `coords` [n][3]; every caller site must be
 * present exactly once.  `time_out` receives the record's time step. */
int hlbx_read_distributions(const char* path, int64_t n_sites, const int32_t* coords,
                            int32_t q, int64_t target_time, double* f_out,
                            uint64_t* time_out);

/* Reason for the last failure of an hlbx_ call on this thread. */
const char* hlbx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
