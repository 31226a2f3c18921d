// model.cpp -- read/write of the io model layout (the datasets and order of
// Model::read, src/Model/model.cpp:54-86, and of every component's read()).
#include <algorithm>
#include <numeric>

#include "model.hpp"

namespace magritte {

namespace {
const std::string CELLS = "Geometry/Cells/";
const std::string RAYS = "Geometry/Rays/";
const std::string BOUNDARY = "Geometry/Boundary/";
const std::string TEMPERATURE = "Thermodynamics/Temperature/";
const std::string TURBULENCE = "Thermodynamics/Turbulence/";
const std::string SPECIES = "Chemistry/Species/";
const std::string LSPEC = "Lines/LineProducingSpecies_";
const std::string RADIATION = "Radiation/";

Double2 to_rows(const Vector3d1 &v)
{
    Double2 out(v.size(), Double1(3));
    for (size_t p = 0; p < v.size(); p++) out[p] = {v[p][0], v[p][1], v[p][2]};
    return out;
}
}  // namespace

// src/Model/parameters.cpp:12-61: every attribute is optional
void Parameters::read(const Io &io)
{
    size_t n;
    if (io.read_number(".ncells", n) == 0) set_ncells(n);
    if (io.read_number(".nrays", n) == 0) set_nrays(n);
    if (io.read_number(".nrays_red", n) == 0) set_nrays_red(n);
    if (io.read_number(".order_min", n) == 0) set_order_min(n);
    if (io.read_number(".order_max", n) == 0) set_order_max(n);
    if (io.read_number(".nboundary", n) == 0) set_nboundary(n);
    if (io.read_number(".nfreqs", n) == 0) set_nfreqs(n);
    if (io.read_number(".nfreqs_red", n) == 0) set_nfreqs_red(n);
    if (io.read_number(".nspecs", n) == 0) set_nspecs(n);
    if (io.read_number(".nlspecs", n) == 0) set_nlspecs(n);
    if (io.read_number(".nlines", n) == 0) set_nlines(n);
    if (io.read_number(".nquads", n) == 0) set_nquads(n);
    double d;
    if (io.read_number(".pop_prec", d) == 0) set_pop_prec(d);
    bool b;
    if (io.read_bool(".use_scattering", b) == 0) set_use_scattering(b);
    if (io.read_bool(".use_Ng_acceleration", b) == 0) set_use_Ng_acceleration(b);
    if (io.read_bool(".spherical_symmetry", b) == 0) set_spherical_symmetry(b);
    if (io.read_bool(".adaptive_ray_tracing", b) == 0) set_adaptive_ray_tracing(b);
}

void Parameters::write(const Io &io) const
{
#define W(name) if (name##_.is_set) io.write_number("." #name, name##_.value);
    W(ncells) W(nrays) W(nrays_red) W(order_min) W(order_max) W(nboundary) W(nfreqs)
    W(nfreqs_red) W(nspecs) W(nlspecs) W(nlines) W(nquads) W(pop_prec)
#undef W
#define WB(name) if (name##_.is_set) io.write_bool("." #name, name##_.value);
    WB(use_scattering) WB(use_Ng_acceleration) WB(spherical_symmetry) WB(adaptive_ray_tracing)
#undef WB
}

// Cells::read (src/Model/Geometry/Cells/cells.cpp:20-93)
void Cells::read(const Io &io, Parameters &parameters)
{
    size_t ncells;
    io.read_length(CELLS + "position", ncells);
    parameters.set_ncells(ncells);
    Double2 pos(ncells, Double1(3)), vel(ncells, Double1(3));
    io.read_array(CELLS + "position", pos);
    io.read_array(CELLS + "velocity", vel);
    position.resize(ncells);
    velocity.resize(ncells);
    for (size_t p = 0; p < ncells; p++) {
        position[p] = {pos[p][0], pos[p][1], pos[p][2]};
        velocity[p] = {vel[p][0], vel[p][1], vel[p][2]};
    }
    n_neighbors.assign(ncells, 0);
    io.read_list(CELLS + "n_neighbors", n_neighbors);
    const size_t tot = std::accumulate(n_neighbors.begin(), n_neighbors.end(), (size_t)0);
    Long1 lin(tot);
    io.read_list(CELLS + "neighbors", lin);
    neighbors.assign(ncells, Long1());
    size_t at = 0;
    for (size_t p = 0; p < ncells; p++) {
        neighbors[p].assign(lin.begin() + at, lin.begin() + at + n_neighbors[p]);
        at += n_neighbors[p];
    }
}

void Cells::write(const Io &io) const
{
    io.write_array(CELLS + "position", to_rows(position));
    io.write_array(CELLS + "velocity", to_rows(velocity));
    Long1 nn(neighbors.size()), lin;
    for (size_t p = 0; p < neighbors.size(); p++) {
        nn[p] = (long)neighbors[p].size();
        lin.insert(lin.end(), neighbors[p].begin(), neighbors[p].end());
    }
    io.write_list(CELLS + "n_neighbors", nn);
    io.write_list(CELLS + "neighbors", lin);
}

// Rays::read / setup_antipodal_rays (src/Model/Geometry/Rays/rays.cpp:22-103)
void Rays::read(const Io &io, Parameters &parameters)
{
    io.read_length(RAYS + "rays", nrays);
    parameters.set_nrays(nrays);
    Double2 arr(nrays, Double1(3));
    io.read_array(RAYS + "rays", arr);
    rays.resize(nrays);
    for (size_t r = 0; r < nrays; r++) rays[r] = {arr[r][0], arr[r][1], arr[r][2]};
    weights.assign(nrays, 0.0);
    io.read_list(RAYS + "weights", weights);
    setup();
}

void Rays::setup()
{
    nrays = rays.size();
    half_nrays = nrays / 2;
    antipod.assign(nrays, 0);
    for (size_t r1 = 0; r1 < nrays; r1++)
        for (size_t r2 = 0; r2 < nrays; r2++) {
            const double x = rays[r1][0] + rays[r2][0];
            const double y = rays[r1][1] + rays[r2][1];
            const double z = rays[r1][2] + rays[r2][2];
            if (((x * x + y * y) + z * z) < 1.0E-9) antipod[r1] = r2;
        }
}

void Rays::write(const Io &io) const
{
    io.write_array(RAYS + "rays", to_rows(rays));
    io.write_list(RAYS + "weights", weights);
}

// Boundary::read (src/Model/Geometry/Boundary/boundary.cpp:19-67), including
// its string parsing of the boundary conditions (each comparison tests for a
// mismatch, so "cmb" ends as Thermal and "zero"/"thermal" as CMB).
void Boundary::read(const Io &io, Parameters &parameters)
{
    const size_t ncells = parameters.ncells();
    cell2boundary_nr.assign(ncells, (long)ncells);
    boundary.assign(ncells, false);
    size_t nboundary;
    io.read_length(BOUNDARY + "boundary2cell_nr", nboundary);
    boundary2cell_nr.assign(nboundary, 0);
    io.read_list(BOUNDARY + "boundary2cell_nr", boundary2cell_nr);
    parameters.set_nboundary(nboundary);
    for (size_t b = 0; b < nboundary; b++) {
        cell2boundary_nr[boundary2cell_nr[b]] = (long)b;
        boundary[boundary2cell_nr[b]] = true;
    }
    boundary_condition.assign(nboundary, Zero);
    boundary_temperature.assign(nboundary, 0.0);
    String1 bc(nboundary);
    io.read_list(BOUNDARY + "boundary_temperature", boundary_temperature);
    io.read_list(BOUNDARY + "boundary_condition", bc);
    for (size_t b = 0; b < nboundary && b < bc.size(); b++) {
        if (bc[b].compare("zero")) boundary_condition[b] = Zero;
        if (bc[b].compare("thermal")) boundary_condition[b] = Thermal;
        if (bc[b].compare("cmb")) boundary_condition[b] = CMB;
    }
}

void Boundary::write(const Io &io) const
{
    io.write_list(BOUNDARY + "boundary2cell_nr", boundary2cell_nr);
    io.write_list(BOUNDARY + "boundary_temperature", boundary_temperature);
    String1 bc;
    for (auto c : boundary_condition) bc.push_back(c == Zero ? "zero" : c == Thermal ? "thermal" : "cmb");
    io.write_list(BOUNDARY + "boundary_condition", bc);
}

void Geometry::read(const Io &io, Parameters &parameters)
{
    cells.read(io, parameters);
    rays.read(io, parameters);
    boundary.read(io, parameters);
}

void Geometry::write(const Io &io) const
{
    cells.write(io);
    rays.write(io);
    boundary.write(io);
}

void Thermodynamics::read(const Io &io, Parameters &parameters)
{
    size_t n;
    io.read_length(TEMPERATURE + "gas", n);
    parameters.set_ncells(n);
    temperature.gas.assign(n, 0.0);
    io.read_list(TEMPERATURE + "gas", temperature.gas);
    io.read_length(TURBULENCE + "vturb2", n);
    parameters.set_ncells(n);
    turbulence.vturb2.assign(n, 0.0);
    io.read_list(TURBULENCE + "vturb2", turbulence.vturb2);
}

void Thermodynamics::write(const Io &io) const
{
    io.write_list(TEMPERATURE + "gas", temperature.gas);
    io.write_list(TURBULENCE + "vturb2", turbulence.vturb2);
}

// Species::read (src/Model/Chemistry/Species/species.cpp:19-56)
void Species::read(const Io &io, Parameters &parameters)
{
    size_t nspecs, ncells;
    io.read_length(SPECIES + "species", nspecs);
    io.read_length(SPECIES + "abundance", ncells);
    parameters.set_ncells(ncells);
    parameters.set_nspecs(nspecs);
    sym.assign(nspecs, "");
    io.read_list(SPECIES + "species", sym);
    abundance.assign(ncells, Double1(nspecs, 0.0));
    io.read_array(SPECIES + "abundance", abundance);
}

void Species::write(const Io &io) const
{
    io.write_list(SPECIES + "species", sym);
    io.write_array(SPECIES + "abundance", abundance);
}

// CollisionPartner::read (CollisionPartner/collisionPartner.cpp:20-60)
void CollisionPartner::read(const Io &io, long l, long c)
{
    const std::string p = LSPEC + std::to_string(l) + "/Linedata/CollisionPartner_" + std::to_string(c) + "/";
    io.read_number(p + ".num_col_partner", num_col_partner);
    io.read_word(p + ".orth_or_para_H2", orth_or_para_H2);
    io.read_length(p + "tmp", ntmp);
    io.read_length(p + "icol", ncol);
    tmp.assign(ntmp, 0.0);
    io.read_list(p + "tmp", tmp);
    icol.assign(ncol, 0);
    jcol.assign(ncol, 0);
    io.read_list(p + "icol", icol);
    io.read_list(p + "jcol", jcol);
    Ce.assign(ntmp, Double1(ncol, 0.0));
    Cd.assign(ntmp, Double1(ncol, 0.0));
    io.read_array(p + "Ce", Ce);
    io.read_array(p + "Cd", Cd);
}

void CollisionPartner::write(const Io &io, long l, long c) const
{
    const std::string p = LSPEC + std::to_string(l) + "/Linedata/CollisionPartner_" + std::to_string(c) + "/";
    io.write_number(p + ".num_col_partner", num_col_partner);
    io.write_word(p + ".orth_or_para_H2", orth_or_para_H2);
    io.write_list(p + "icol", icol);
    io.write_list(p + "jcol", jcol);
    io.write_list(p + "tmp", tmp);
    io.write_array(p + "Ce", Ce);
    io.write_array(p + "Cd", Cd);
}

// Linedata::read (Linedata/linedata.cpp:25-88)
void Linedata::read(const Io &io, long l)
{
    const std::string p = LSPEC + std::to_string(l) + "/Linedata/";
    io.read_number(p + ".num", num);
    io.read_word(p + ".sym", sym);
    io.read_number(p + ".inverse_mass", inverse_mass);
    io.read_number(p + ".nlev", nlev);
    io.read_number(p + ".nrad", nrad);
    irad.assign(nrad, 0); jrad.assign(nrad, 0);
    io.read_list(p + "irad", irad);
    io.read_list(p + "jrad", jrad);
    energy.assign(nlev, 0.0); weight.assign(nlev, 0.0);
    io.read_list(p + "energy", energy);
    io.read_list(p + "weight", weight);
    frequency.assign(nrad, 0.0); A.assign(nrad, 0.0); Bs.assign(nrad, 0.0); Ba.assign(nrad, 0.0);
    io.read_list(p + "frequency", frequency);
    io.read_list(p + "A", A);
    io.read_list(p + "Bs", Bs);
    io.read_list(p + "Ba", Ba);
    io.read_length(p + "CollisionPartner_", ncolpar);
    colpar.resize(ncolpar);
    for (size_t c = 0; c < ncolpar; c++) colpar[c].read(io, l, (long)c);
}

void Linedata::write(const Io &io, long l) const
{
    const std::string p = LSPEC + std::to_string(l) + "/Linedata/";
    io.write_number(p + ".num", num);
    io.write_word(p + ".sym", sym);
    io.write_number(p + ".inverse_mass", inverse_mass);
    io.write_number(p + ".nlev", nlev);
    io.write_number(p + ".nrad", nrad);
    io.write_list(p + "irad", irad);
    io.write_list(p + "jrad", jrad);
    io.write_list(p + "energy", energy);
    io.write_list(p + "weight", weight);
    io.write_list(p + "frequency", frequency);
    io.write_list(p + "A", A);
    io.write_list(p + "Bs", Bs);
    io.write_list(p + "Ba", Ba);
    for (size_t c = 0; c < colpar.size(); c++) colpar[c].write(io, l, (long)c);
}

void Quadrature::read(const Io &io, long l, Parameters &parameters)
{
    const std::string p = LSPEC + std::to_string(l) + "/Quadrature/";
    size_t nquads;
    io.read_length(p + "weights", nquads);
    parameters.set_nquads(nquads);
    weights.assign(nquads, 0.0);
    roots.assign(nquads, 0.0);
    io.read_list(p + "weights", weights);
    io.read_list(p + "roots", roots);
}

void Quadrature::write(const Io &io, long l) const
{
    const std::string p = LSPEC + std::to_string(l) + "/Quadrature/";
    io.write_list(p + "weights", weights);
    io.write_list(p + "roots", roots);
}

// LineProducingSpecies::read / write (lineProducingSpecies.cpp:21-206)
void LineProducingSpecies::read(const Io &io, long l, Parameters &parameters)
{
    linedata.read(io, l);
    quadrature.read(io, l, parameters);
    ncells = (long)parameters.ncells();
    const long V = linedata.nlev, K = linedata.nrad;
    const size_t Q = parameters.nquads();
    Jeff.assign(ncells, Double1(K, 0.0));
    Jlin.assign(ncells, Double1(K, 0.0));
    Jdif.assign(ncells, Double1(K, 0.0));
    nr_line.assign(ncells, Long2(K, Long1(Q, 0)));
    lambda.Ls.assign(ncells, Double2(K));
    lambda.nr.assign(ncells, Long2(K));
    lambda.size.assign(ncells * K, 0);
    population.assign(ncells * V, 0.0);
    population_prev1.assign(ncells * V, 0.0);
    population_prev2.assign(ncells * V, 0.0);
    population_prev3.assign(ncells * V, 0.0);
    population_tot.assign(ncells, 0.0);
    const std::string p = LSPEC + std::to_string(l) + "/";
    io.read_list(p + "population_tot", population_tot);
    read_populations(io, l, "");
    Double2 prev(ncells, Double1(V));
    Double1 *dst[3] = {&population_prev1, &population_prev2, &population_prev3};
    const char *names[3] = {"population_prev1", "population_prev2", "population_prev3"};
    for (int h = 0; h < 3; h++)
        if (io.read_array(p + names[h], prev) == 0)
            for (long q = 0; q < ncells; q++)
                for (long i = 0; i < V; i++) (*dst[h])[index(q, i)] = prev[q][i];
}

int LineProducingSpecies::read_populations(const Io &io, long l, const std::string &tag)
{
    const std::string p = LSPEC + std::to_string(l) + "/";
    const long V = linedata.nlev, K = linedata.nrad;
    Double2 pops(ncells, Double1(V));
    const int err = io.read_array(p + "population" + tag, pops);
    if (err == 0)
        for (long q = 0; q < ncells; q++)
            for (long i = 0; i < V; i++) population[index(q, i)] = pops[q][i];
    Double2 jl(ncells, Double1(K)), je(ncells, Double1(K));
    if (io.read_array(p + "J_lin" + tag, jl) == 0) Jlin = jl;
    if (io.read_array(p + "J_eff" + tag, je) == 0) Jeff = je;
    return err;
}

int LineProducingSpecies::write_populations(const Io &io, long l, const std::string &tag) const
{
    const std::string p = LSPEC + std::to_string(l) + "/";
    const long V = linedata.nlev;
    if (population.size() != (size_t)(ncells * V))
        throw std::runtime_error("write_populations: population holds " + std::to_string(population.size()) +
                                 " values, ncells * nlev = " + std::to_string(ncells * V));
    Double2 pops(ncells, Double1(V));
    for (long q = 0; q < ncells; q++)
        for (long i = 0; i < V; i++) pops[q][i] = population[index(q, i)];
    io.write_array(p + "population" + tag, pops);
    io.write_array(p + "J_lin" + tag, Jlin);
    io.write_array(p + "J_eff" + tag, Jeff);
    return 0;
}

void LineProducingSpecies::write(const Io &io, long l) const
{
    linedata.write(io, l);
    quadrature.write(io, l);
    write_populations(io, l, "");
    const std::string p = LSPEC + std::to_string(l) + "/";
    io.write_list(p + "population_tot", population_tot);
    const long V = linedata.nlev;
    const Double1 *src[3] = {&population_prev1, &population_prev2, &population_prev3};
    const char *names[3] = {"population_prev1", "population_prev2", "population_prev3"};
    for (int h = 0; h < 3; h++) {
        if (src[h]->size() != (size_t)(ncells * V))
            throw std::runtime_error(std::string("write: ") + names[h] + " does not hold ncells * nlev values");
        Double2 prev(ncells, Double1(V));
        for (long q = 0; q < ncells; q++)
            for (long i = 0; i < V; i++) prev[q][i] = (*src[h])[index(q, i)];
        io.write_array(p + names[h], prev);
    }
}

// Lines::read / setup (src/Model/Lines/lines.cpp:22-87)
void Lines::read(const Io &io, Parameters &parameters)
{
    size_t nlspecs;
    io.read_length(LSPEC, nlspecs);
    parameters.set_nlspecs(nlspecs);
    lineProducingSpecies.resize(nlspecs);
    for (size_t l = 0; l < nlspecs; l++) lineProducingSpecies[l].read(io, (long)l, parameters);
    size_t nlines = 0;
    for (auto &ls : lineProducingSpecies) nlines += ls.linedata.nrad;
    parameters.set_nlines(nlines);
    line.clear();
    line_index.clear();
    for (auto &ls : lineProducingSpecies)
        for (long k = 0; k < ls.linedata.nrad; k++) {
            line.push_back(ls.linedata.frequency[k]);
            line_index.push_back((long)line_index.size());
        }
    std::vector<size_t> order(line.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return line[a] < line[b]; });
    Double1 l2(line.size());
    Long1 i2(line.size());
    for (size_t q = 0; q < order.size(); q++) { l2[q] = line[order[q]]; i2[q] = line_index[order[q]]; }
    line = l2;
    line_index = i2;
    emissivity.assign(parameters.ncells() * nlines, 0.0);
    opacity.assign(parameters.ncells() * nlines, 0.0);
}

void Lines::write(const Io &io) const
{
    for (size_t l = 0; l < lineProducingSpecies.size(); l++) lineProducingSpecies[l].write(io, (long)l);
}

// Radiation::read (src/Model/Radiation/radiation.cpp:94-135) with the
// frequency counts of Frequencies::read (frequencies.cpp:24-80)
void Radiation::read(const Io &io, Parameters &parameters)
{
    const size_t nfreqs = parameters.nlines() * parameters.nquads();
    parameters.set_nfreqs(nfreqs);
    parameters.set_nfreqs_red(nfreqs);
    const size_t ncells = parameters.ncells();
    const size_t nrays_red = parameters.nrays() / 2;
    parameters.set_nrays_red(nrays_red);
    frequencies.nu.assign(ncells, Double1(nfreqs, 0.0));
    frequencies.appears_in_line_integral.assign(nfreqs, false);
    frequencies.corresponding_l_for_spec.assign(nfreqs, (long)nfreqs);
    frequencies.corresponding_k_for_tran.assign(nfreqs, (long)nfreqs);
    frequencies.corresponding_z_for_line.assign(nfreqs, (long)nfreqs);
    J.assign(ncells * nfreqs, 0.0);
    I_bdy.assign(nrays_red, Double2(parameters.nboundary(), Double1(nfreqs, 0.0)));
    // u of every reduced ray with scattering (radiation.cpp:117-126)
    if (parameters.use_scattering_.is_set && parameters.use_scattering())
        u.assign(nrays_red, Double1(ncells * nfreqs, 0.0));
    else
        u.clear();
    (void)io;
}

void Frequencies::write(const Io &io) const { io.write_array(RADIATION + "Frequencies/nu", nu); }

void Radiation::write(const Io &io) const
{
    frequencies.write(io);
    if (!u.empty()) io.write_array(RADIATION + "u", u);
}

void Model::read(const Io &io)
{
    parameters.read(io);
    geometry.read(io, parameters);
    thermodynamics.read(io, parameters);
    chemistry.read(io, parameters);
    lines.read(io, parameters);
    radiation.read(io, parameters);
}

void Model::write(const Io &io) const
{
    parameters.write(io);
    geometry.write(io);
    thermodynamics.write(io);
    chemistry.write(io);
    lines.write(io);
    radiation.write(io);
}

}  // namespace magritte
