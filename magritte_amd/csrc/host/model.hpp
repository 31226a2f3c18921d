// model.hpp -- Model aggregate of the drop-in host layer, with the member
// names and io layout of the reference (src/Model/model.hpp:23-44 and the
// read()/write() of every component).  Vectors are std::vector; positions and
// velocities are {x, y, z} triples.
#pragma once

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "io.hpp"

namespace magritte {

using Vector3d = std::array<double, 3>;
using Vector3d1 = std::vector<Vector3d>;

// SetOnce (src/Tools/setOnce.hpp:15-85): a value may be set again only with
// the same value; reading an unset value throws.
template <typename T>
struct SetOnce {
    bool is_set = false;
    T value{};
    void set(const T &v)
    {
        if (is_set && !(v == value)) throw std::runtime_error("DoubleSetException: parameter set twice with different values");
        value = v;
        is_set = true;
    }
    T get() const
    {
        if (!is_set) throw std::runtime_error("GetBeforeSetException: parameter read before it was set");
        return value;
    }
};

struct Parameters {
    SetOnce<size_t> ncells_, nrays_, nrays_red_, order_min_, order_max_, nboundary_, nfreqs_,
        nfreqs_red_, nspecs_, nlspecs_, nlines_, nquads_;
    SetOnce<double> pop_prec_;
    SetOnce<bool> use_scattering_, use_Ng_acceleration_, spherical_symmetry_, adaptive_ray_tracing_;
    long   n_off_diag = 0;            // parameters.hpp:87
    double max_width_fraction = 0.5;  // parameters.hpp:89

#define MGR_PARAM(T, name) \
    void set_##name(const T &v) { name##_.set(v); } \
    T name() const { return name##_.get(); }
    MGR_PARAM(size_t, ncells)
    MGR_PARAM(size_t, nrays)
    MGR_PARAM(size_t, nrays_red)
    MGR_PARAM(size_t, order_min)
    MGR_PARAM(size_t, order_max)
    MGR_PARAM(size_t, nboundary)
    MGR_PARAM(size_t, nfreqs)
    MGR_PARAM(size_t, nfreqs_red)
    MGR_PARAM(size_t, nspecs)
    MGR_PARAM(size_t, nlspecs)
    MGR_PARAM(size_t, nlines)
    MGR_PARAM(size_t, nquads)
    MGR_PARAM(double, pop_prec)
    MGR_PARAM(bool, use_scattering)
    MGR_PARAM(bool, use_Ng_acceleration)
    MGR_PARAM(bool, spherical_symmetry)
    MGR_PARAM(bool, adaptive_ray_tracing)
#undef MGR_PARAM

    void read(const Io &io);
    void write(const Io &io) const;
};

struct Cells {
    Vector3d1 position, velocity;
    Long1 n_neighbors;
    Long2 neighbors;
    void read(const Io &io, Parameters &parameters);
    void write(const Io &io) const;
};

struct Rays {
    Vector3d1 rays;
    Double1 weights;
    Size1 antipod;
    size_t nrays = 0, half_nrays = 0;
    void read(const Io &io, Parameters &parameters);
    void write(const Io &io) const;
    void setup();
};

enum BoundaryCondition { Zero, Thermal, CMB };

struct Boundary {
    Long1 boundary2cell_nr, cell2boundary_nr;
    Bool1 boundary;
    std::vector<BoundaryCondition> boundary_condition;
    Double1 boundary_temperature;
    void read(const Io &io, Parameters &parameters);
    void write(const Io &io) const;
};

struct Geometry {
    Cells cells;
    Rays rays;
    Boundary boundary;
    long max_npoints_on_rays = -1;
    void read(const Io &io, Parameters &parameters);
    void write(const Io &io) const;
};

struct Temperature { Double1 gas; };
struct Turbulence { Double1 vturb2; };
struct Thermodynamics {
    Temperature temperature;
    Turbulence turbulence;
    void read(const Io &io, Parameters &parameters);
    void write(const Io &io) const;
};

struct Species {
    Double2 abundance;
    String1 sym;
    void read(const Io &io, Parameters &parameters);
    void write(const Io &io) const;
};
struct Chemistry {
    Species species;
    void read(const Io &io, Parameters &parameters) { species.read(io, parameters); }
    void write(const Io &io) const { species.write(io); }
};

struct CollisionPartner {
    long num_col_partner = 0;
    std::string orth_or_para_H2 = "n";
    size_t ntmp = 0, ncol = 0;
    Long1 icol, jcol;
    Double1 tmp;
    Double2 Ce, Cd;
    void read(const Io &io, long l, long c);
    void write(const Io &io, long l, long c) const;
};

struct Linedata {
    long num = 0;
    std::string sym;
    double inverse_mass = 0.0;
    long nlev = 0, nrad = 0;
    Long1 irad, jrad;
    Double1 energy, weight, frequency, A, Ba, Bs;
    size_t ncolpar = 0;
    std::vector<CollisionPartner> colpar;
    void read(const Io &io, long l);
    void write(const Io &io, long l) const;
};

struct Quadrature {
    Double1 roots, weights;
    void read(const Io &io, long l, Parameters &parameters);
    void write(const Io &io, long l) const;
};

// Lambda (src/Model/Lines/LineProducingSpecies/Lambda/lambda.hpp): Ls[p][k]
// and nr[p][k] list the ALO elements of (p, k); with n_off_diag = 0 there is
// one element, at nr = p.
struct Lambda {
    Double3 Ls;
    Long3 nr;
    Size1 size;
    Double1 Lss;
    Long1 nrs;
};

struct LineProducingSpecies {
    Linedata linedata;
    Quadrature quadrature;
    Lambda lambda;
    Double2 Jlin, Jeff, Jdif;
    Long3 nr_line;
    double relative_change_mean = 0.0, relative_change_max = 0.0, fraction_not_converged = 0.0;
    Double1 population, population_tot, population_prev1, population_prev2, population_prev3;
    long ncells = 0;
    long index(long p, long i) const { return i + p * linedata.nlev; }
    void read(const Io &io, long l, Parameters &parameters);
    void write(const Io &io, long l) const;
    int read_populations(const Io &io, long l, const std::string &tag);
    int write_populations(const Io &io, long l, const std::string &tag) const;
};

struct Lines {
    std::vector<LineProducingSpecies> lineProducingSpecies;
    Double1 emissivity, opacity, line;
    Long1 line_index;
    void read(const Io &io, Parameters &parameters);
    void write(const Io &io) const;
};

struct Frequencies {
    Double2 nu;
    Bool1 appears_in_line_integral;
    Long1 corresponding_l_for_spec, corresponding_k_for_tran, corresponding_z_for_line;
    void write(const Io &io) const;
};

struct Radiation {
    Frequencies frequencies;
    Double1 J;         // [ncells * nfreqs], index f + p nfreqs
    Double3 I_bdy;     // [nrays_red][nboundary][nfreqs]
    Double2 u, v;
    long get_index(long p, long f) const { return f + p * (long)(frequencies.nu.empty() ? 0 : frequencies.nu[0].size()); }
    double get_J(long p, long f) const { return J[get_index(p, f)]; }
    void read(const Io &io, Parameters &parameters);
    void write(const Io &io) const;
};

struct Model {
    Parameters parameters;
    Geometry geometry;
    Chemistry chemistry;
    Lines lines;
    Thermodynamics thermodynamics;
    Radiation radiation;
    virtual ~Model() = default;
    void read(const Io &io);
    void write(const Io &io) const;
};

}  // namespace magritte
