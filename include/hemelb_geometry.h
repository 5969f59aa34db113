/*
 * hemelb_geometry.h -- C ABI for geometry input to the MI355X HemeLB timestep.
 *
 * A geometry is what a HemeLB .gmy v4 file holds (Code/io/formats/geometry.h,
 * Code/geometry/GeometryReader.cc:556-653): fluid sites on a block grid, and
 * for every fluid site the 26 Moore-neighbourhood links with their cut type
 * (none / wall / inlet / outlet), iolet index and cut distance (a fraction of
 * the link vector), plus a wall normal for sites with wall links.
 *
 * hlbg_project() turns it into the per-lattice site arrays that
 * geometry::Domain holds (Code/geometry/Domain.cc:69-357): sites in the
 * reference's local order (collision type 0..5, then octree leaf (Morton)
 * order of blocks, then z-fastest order inside a block), wall/iolet link
 * bitmasks (bit d-1 for lattice direction d, SiteDataBare.cc:41-66), one
 * iolet id per site, cut distances per lattice direction and site type.
 *
 * All functions return 0 on success and a negative code on failure;
 * hlbg_last_error() describes the last failure on the calling thread.
 */
#ifndef HEMELB_GEOMETRY_H
#define HEMELB_GEOMETRY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hlbg_geometry hlbg_geometry;

/* Cut types of a link (io/formats/geometry.h:45-51). */
enum { HLBG_CUT_NONE = 0, HLBG_CUT_WALL = 1, HLBG_CUT_INLET = 2, HLBG_CUT_OUTLET = 3 };

/* Site types (geometry::SiteType: FLUID / INLET / OUTLET). */
enum { HLBG_SITE_FLUID = 1, HLBG_SITE_INLET = 2, HLBG_SITE_OUTLET = 3 };

/* Description of one inlet or outlet plane, in lattice units. The normal
 * points into the fluid. */
typedef struct {
  double position[3];
  double normal[3];
  double radius;
} hlbg_iolet;

/* Straight cylinder along z: `radius` and `length` in voxels; fluid sites are
 * those with (x-cx)^2+(y-cy)^2 < radius^2 and z in [pad, pad+length).  The
 * inlet plane sits half a voxel below the first fluid layer and the outlet
 * half a voxel above the last.  Cut distances are exact ray-cylinder and
 * ray-plane intersections (geometry-tool CylinderGenerator.cpp:91-203). */
int hlbg_cylinder(double radius, int32_t length, int32_t block_size,
                  hlbg_geometry** out);

/* Straight cylinder along z with a Gaussian stenosis:
 * R(z) = radius * (1 - delta * exp(-(z-z0)^2 / (2 w^2))), delta chosen so the
 * area reduction at z0 equals `severity`. */
int hlbg_stenosed_cylinder(double radius, int32_t length, double severity,
                           double z0, double width, int32_t block_size,
                           hlbg_geometry** out);

/* Seeded bifurcating vessel tree: `generations` levels of capsules (a root and
 * binary splits), Murray's law r_p^3 = r_1^3 + r_2^3 with split fraction
 * U(0.4,0.6), segment length 8 r U(0.8,1.2), branch half-angle U(25,45) deg
 * and uniform random azimuth.  The fluid is the union of capsules; one inlet
 * at the root, one outlet per leaf.  Cut distances are exact exits from the
 * capsule union or crossings of the iolet planes. */
int hlbg_vessel_tree(uint64_t seed, int32_t generations, double root_radius,
                     int32_t block_size, hlbg_geometry** out);

int hlbg_read_gmy(const char* path, hlbg_geometry** out);
int hlbg_write_gmy(const hlbg_geometry* g, const char* path);

/* Replaces the iolet descriptions (the .gmy file holds none; they come from
 * the simulation configuration). */
int hlbg_set_iolets(hlbg_geometry* g, const hlbg_iolet* inlets, int32_t n_in,
                    const hlbg_iolet* outlets, int32_t n_out);

void hlbg_free(hlbg_geometry* g);

int64_t hlbg_fluid_site_count(const hlbg_geometry* g);
/* Block grid: blocks[3] and block_size. */
int hlbg_block_info(const hlbg_geometry* g, int32_t* blocks, int32_t* block_size);
int32_t hlbg_iolet_count(const hlbg_geometry* g, int32_t outlet);
int hlbg_get_iolet(const hlbg_geometry* g, int32_t outlet, int32_t index,
                   hlbg_iolet* out);

/* Raw 26-link data in .gmy site order: coords [n][3]; link_type [n][26]
 * (uint8); link_iolet [n][26]; link_dist [n][26] (-1 where no cut); normal
 * [n][3] and has_normal [n].  Any output pointer may be NULL. */
int hlbg_get_links(const hlbg_geometry* g, int32_t* coords, uint8_t* link_type,
                   int32_t* link_iolet, float* link_dist, float* normal,
                   uint8_t* has_normal);

/* Lattice projection (Q = 15 or 19) in the reference's local site order.
 * Arrays are caller-owned with n = hlbg_fluid_site_count() entries:
 *   coords [n][3], site_type [n], wall_mask [n], iolet_mask [n], iolet_id [n],
 *   cut [n][Q-1] (float, -1 where the link is not cut), normal [n][3]
 *   (float, zero where no wall normal).  Any output pointer may be NULL. */
int hlbg_project(const hlbg_geometry* g, int32_t Q, int32_t* coords,
                 uint32_t* site_type, uint32_t* wall_mask, uint32_t* iolet_mask,
                 int32_t* iolet_id, float* cut, float* normal);

/* Signed distance field of the crop [origin, origin + dims) (voxel centres,
 * out[(x * dims[1] + y) * dims[2] + z]): at a fluid site the Euclidean
 * distance to the nearest non-fluid voxel, elsewhere minus the distance to
 * the nearest fluid site (voxel units; a fully solid or fully fluid crop
 * gives +-1e10).  The input of the learned warm start; no counterpart in
 * the reference. */
int hlbg_signed_distance(const hlbg_geometry* g, const int32_t origin[3],
                         const int32_t dims[3], float* out);

const char* hlbg_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
