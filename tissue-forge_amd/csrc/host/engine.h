// Host engine: the model objects of one simulated system and the device
// toggle.  Mirrors the parts of the reference's `struct engine` / `struct
// space` / Universe that the engine_step hot path reads
// (source/mdcore/include/tfEngine.h:164-418, tfSpace.h:127-243).
//
// The host engine owns every object.  Stepping is done only by the device
// module: the reference's own CPU engine is not reimplemented here.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "potential.h"

namespace tfg {

struct DeviceEnsemble;   // device module (device.hip)

// Particle flags (tfParticle.h:69-77).
enum ParticleFlags : uint32_t {
    PFLAG_GHOST = 1u << 0,
    PFLAG_CLUSTER = 1u << 1,
    PFLAG_BOUND = 1u << 2,
    PFLAG_FROZEN_X = 1u << 3,
    PFLAG_FROZEN_Y = 1u << 4,
    PFLAG_FROZEN_Z = 1u << 5,
};

enum Dynamics : int { DYN_NEWTONIAN = 0, DYN_OVERDAMPED = 1 };   // tfParticle.h:63-64
enum FluxKind : int { FLUX_FICK = 0, FLUX_SECRETE = 1, FLUX_UPTAKE = 2 };   // tfFlux.h:40-42
enum DpdNoise : int { DPD_NOISE_UNIT = 0, DPD_NOISE_PHILOX = 1 };
// EngineIntegrator (tfEngine.h:89-92), chosen by the Universe configuration
// (tfSimulator.cpp:218-229).
enum Integrator : int { INTEG_FORWARD_EULER = 0, INTEG_RUNGE_KUTTA_4 = 1 };
// How a nonbonded sum potential is evaluated: the constituents in order up to
// the first one in range (the CPU engine: potential_eval_super_ex evaluates
// `pca || pcb`, tf_potential_eval.h:211-215), or every constituent (the
// reference's GPU path, tfRunner_cuda.cu:436-560, and Potential::operator()).
enum SumRule : int { SUM_FIRST_IN_RANGE = 0, SUM_ALL = 1 };

// BoundaryConditionKind (tfBoundaryConditions.h:38-46).
enum BoundaryKind : unsigned {
    BC_VELOCITY = 1u << 0,
    BC_PERIODIC = 1u << 1,
    BC_FREESLIP = 1u << 2,
    BC_POTENTIAL = 1u << 3,
    BC_NO_SLIP = 1u << 4,
    BC_RESETTING = 1u << 5,
    BC_ACTIVE = BC_FREESLIP | BC_VELOCITY | BC_POTENTIAL,
};
// Faces, low and high side of each axis: left/right (x), front/back (y),
// bottom/top (z).
enum Face : int { FACE_LEFT = 0, FACE_RIGHT = 1, FACE_FRONT = 2, FACE_BACK = 3, FACE_BOTTOM = 4, FACE_TOP = 5 };

// One face's BoundaryCondition (tfBoundaryConditions.h:48-78): kind, wall
// velocity, restore coefficient and the wall's particle radius.
struct BoundaryFace {
    unsigned kind = BC_PERIODIC;
    double vel[3] = {0, 0, 0};
    double restore = 1.0;
    double radius = 0.0;
};

struct ParticleType {
    std::string name;
    double mass = 1.0, imass = 1.0, radius = 1.0;
    int dynamics = DYN_NEWTONIAN;
};

struct FluxEntry {
    int kind = FLUX_FICK;
    int type_a = 0, type_b = 0;
    int index_a = 0, index_b = 0;
    double coef = 0, decay = 0, target = 0, cutoff = 0;
};

struct BondRec { int32_t i, j; int pot; };
struct AngleRec { int32_t i, j, k; int pot; };

// Cell-grid geometry, computed exactly as space_init / engine_init do
// (tfEngine.cpp:1722-1727, tfSpace.cpp:737-770, 893-999).
struct Geometry {
    double origin[3] = {0, 0, 0};
    double dim[3] = {10, 10, 10};
    int cells[3] = {1, 1, 1};   // as requested; cdim can be one less (floor quirk)
    int cdim[3] = {1, 1, 1};
    double h[3] = {10, 10, 10};
    double ih[3] = {0.1, 0.1, 0.1};
    int span[3] = {1, 1, 1};
    double cutoff = 1.0, cutoff2 = 1.0;
    int ncells = 1;
    // Neighbour-cell offsets (dx, dy, dz) within span whose closest distance
    // is inside the cutoff (the task-building cull of tfSpace.cpp:949-954).
    std::vector<int> stencil;   // 3 ints per offset, includes (0, 0, 0)
};

int make_geometry(Geometry &g, const double origin[3], const double dim[3],
                  const int cells[3], double cutoff);

class Engine;
// Engine state as a JSON document and back (checkpoint.cpp).  engine_from_json
// fills a fresh (default-constructed) engine.
std::string engine_to_json(const Engine &e);
int engine_from_json(Engine &e, const char *text, size_t len);

// Per-particle host record, indexed by particle id.
struct HostParticle {
    int32_t type = 0;
    uint32_t flags = 0;
    int32_t cid = 0;          // cell index (i*cdim1 + j)*cdim2 + k
    double x[3] = {0, 0, 0};  // relative to the cell origin
    double v[3] = {0, 0, 0};
    double f[4] = {0, 0, 0, 0};   // force and number density of the last evaluation
    double imass = 1.0, radius = 1.0;
};

struct Energies {
    double epot = 0, epot_nonbond = 0, epot_bond = 0, epot_angle = 0, ekin = 0, virial = 0;
};

class Engine {
  public:
    Geometry geom;
    std::vector<ParticleType> types;
    int max_types = 64;
    std::vector<std::shared_ptr<Potential>> pots;   // registry, index = pot id
    std::vector<int> potmap;                        // max_types^2 -> pot id or -1
    std::vector<FluxEntry> fluxes;
    int nspecies = 0;
    std::vector<HostParticle> parts;
    std::vector<double> species;   // nspecies per particle
    std::vector<double> species_q; // pending flux per particle
    std::vector<BondRec> bonds;
    std::vector<AngleRec> angles;

    double dt = 0.01;
    int nr_fluxsteps = 1;
    double dt_flux = 0.01;
    double max_dist_fraction = 0.05;   // tfEngine.cpp:1865
    int integrator = INTEG_FORWARD_EULER;
    int sum_rule = SUM_FIRST_IN_RANGE;
    BoundaryFace bc_set[6];        // faces as set (every face periodic by default)
    // Neighbour lists with a skin (0: off); capacity in entries per particle
    // (0: from the density).  The pair set of every step is unchanged.
    double verlet_skin = 0.0;
    int verlet_capacity = 0;
    std::vector<int> wallpot;      // 6 * max_types: pot id of (face, type) or -1
    // Launch settings applied at the next upload (cuda::EngineConfig::
    // setBlocks / setThreads, tfEngineConfig.cpp:98-124); 0 keeps the default.
    int launch_blocks = 0, launch_threads = 0;
    int64_t time = 0;                  // step counter (engine->time)
    int dpd_noise = DPD_NOISE_UNIT;
    uint64_t dpd_seed = 0;

    Energies energies;
    DeviceEnsemble *device = nullptr;   // non-null while on the device
    int device_index = -1;              // position inside that ensemble

    int init(const double origin[3], const double dim[3], const int cells[3],
             double cutoff, int max_types, double dt, int nr_fluxsteps);
    int add_type(const std::string &name, double mass, double radius, int dynamics);
    // Registers a copy of p; an identical table already registered is reused,
    // so a potential bound to several type pairs keeps a single table.
    int add_potential(const Potential &p);
    int bind(int pot, int ti, int tj);
    int add_flux(int kind, int ta, int tb, int ia, int ib, double k, double decay,
                 double target, double cutoff);
    // BoundaryCondition::init (tfBoundaryConditions.cpp:151-171).
    int set_boundary(int face, unsigned kind, const double vel[3], double restore, double radius);
    // BoundaryCondition::set_potential; face -1: all six faces.
    int set_wall_potential(int face, int type, int pot);
    // The faces in effect: check_periodicy (tfBoundaryConditions.cpp:192-217)
    // makes both faces of an axis periodic when one is.  Returns the
    // periodic axes as bits.
    unsigned boundaries(BoundaryFace out[6]) const;
    int add_particles(int n, const int *types, const double *x, const double *v,
                      const double *sp, int32_t *ids_out);
    int add_bonds(int pot, int n, const int32_t *ij);
    int add_angles(int pot, int n, const int32_t *ijk);
};

}  // namespace tfg
