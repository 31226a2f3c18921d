// simulation.cpp -- bodies of the drop-in Simulation methods over the C-ABI.
#include "simulation.hpp"

#include <algorithm>
#include <cstdio>
#include <stdexcept>

namespace magritte {

Simulation::~Simulation() { reset_device(); }

void Simulation::reset_device()
{
    if (ctx_) mgr_destroy(ctx_);
    ctx_ = nullptr;
}

void Simulation::check(int rc, const char *what)
{
    if (rc != MGR_OK) throw std::runtime_error(std::string(what) + " failed: " + mgr_last_error(ctx_));
}

// Upload of the Model (the arrays Model::read filled) into a device context.
void Simulation::ensure_context()
{
    if (ctx_) return;
    const size_t S = lines.lineProducingSpecies.size();
    if (S < 1) throw std::runtime_error("the model has no line-producing species");
    const Cells &cells = geometry.cells;
    const size_t N = cells.position.size();
    const size_t R = geometry.rays.rays.size();
    auto need = [](bool ok, const char *what) {
        if (!ok) throw std::runtime_error(std::string("model member has the wrong size: ") + what);
    };
    need(cells.velocity.size() == N, "geometry.cells.velocity");
    need(cells.neighbors.size() == N, "geometry.cells.neighbors");
    need(geometry.rays.weights.size() == R, "geometry.rays.weights");
    need(thermodynamics.temperature.gas.size() == N, "thermodynamics.temperature.gas");
    need(thermodynamics.turbulence.vturb2.size() == N, "thermodynamics.turbulence.vturb2");
    need(chemistry.species.abundance.size() == N, "chemistry.species.abundance");
    const size_t nq = lines.lineProducingSpecies[0].quadrature.roots.size();
    for (const LineProducingSpecies &ls : lines.lineProducingSpecies) {
        const Linedata &ld = ls.linedata;
        for (const auto &row : chemistry.species.abundance)
            need(row.size() == chemistry.species.abundance[0].size() && (long)row.size() > ld.num,
                 "chemistry.species.abundance rows");
        need(ld.irad.size() == (size_t)ld.nrad && ld.jrad.size() == (size_t)ld.nrad &&
                 ld.A.size() == (size_t)ld.nrad && ld.Bs.size() == (size_t)ld.nrad &&
                 ld.Ba.size() == (size_t)ld.nrad && ld.frequency.size() == (size_t)ld.nrad,
             "linedata radiative arrays (nrad)");
        need(ld.energy.size() == (size_t)ld.nlev && ld.weight.size() == (size_t)ld.nlev,
             "linedata level arrays (nlev)");
        need(ls.quadrature.weights.size() == ls.quadrature.roots.size() && ls.quadrature.roots.size() == nq,
             "quadrature (nquads, the same for every species)");
        for (const auto &cp : ld.colpar) {
            need(cp.jcol.size() == cp.icol.size(), "collision partner jcol");
            need(cp.Cd.size() == cp.tmp.size() && cp.Ce.size() == cp.tmp.size(), "collision partner Cd/Ce rows");
            for (size_t t = 0; t < cp.tmp.size(); t++)
                need(cp.Cd[t].size() == cp.icol.size() && cp.Ce[t].size() == cp.icol.size(),
                     "collision partner Cd/Ce");
        }
    }

    Double1 pos(3 * N), vel(3 * N), rays(3 * R), abundance;
    Long1 n_nb(N), nbs;
    for (size_t p = 0; p < N; p++)
        for (int d = 0; d < 3; d++) {
            pos[3 * p + d] = cells.position[p][d];
            vel[3 * p + d] = cells.velocity[p][d];
        }
    for (size_t p = 0; p < N; p++) {
        n_nb[p] = (long)cells.neighbors[p].size();
        nbs.insert(nbs.end(), cells.neighbors[p].begin(), cells.neighbors[p].end());
    }
    for (size_t r = 0; r < R; r++)
        for (int d = 0; d < 3; d++) rays[3 * r + d] = geometry.rays.rays[r][d];
    const auto &ab = chemistry.species.abundance;
    const size_t nspecs = ab.empty() ? 0 : ab[0].size();
    for (size_t p = 0; p < N; p++) abundance.insert(abundance.end(), ab[p].begin(), ab[p].end());

    std::vector<int64_t> n_nb64(n_nb.begin(), n_nb.end()), nbs64(nbs.begin(), nbs.end());
    std::vector<int64_t> b2c(geometry.boundary.boundary2cell_nr.begin(), geometry.boundary.boundary2cell_nr.end());

    // every species' line data, flattened as the C-ABI takes it
    struct SpeciesArrays {
        std::vector<int64_t> irad, jrad, cp_num, cp_ncol, cp_ntmp, icol, jcol;
        std::string cp_op;
        Double1 tmp, Cd, Ce;
    };
    std::vector<SpeciesArrays> sa(S);
    std::vector<mgr_species_view> sv(S);
    for (size_t l = 0; l < S; l++) {
        const LineProducingSpecies &ls = lines.lineProducingSpecies[l];
        const Linedata &ld = ls.linedata;
        SpeciesArrays &a = sa[l];
        a.irad.assign(ld.irad.begin(), ld.irad.end());
        a.jrad.assign(ld.jrad.begin(), ld.jrad.end());
        for (const auto &cp : ld.colpar) {
            a.cp_num.push_back(cp.num_col_partner);
            a.cp_ncol.push_back((int64_t)cp.icol.size());
            a.cp_ntmp.push_back((int64_t)cp.tmp.size());
            a.cp_op.push_back(cp.orth_or_para_H2.empty() ? 'n' : cp.orth_or_para_H2[0]);
            a.icol.insert(a.icol.end(), cp.icol.begin(), cp.icol.end());
            a.jcol.insert(a.jcol.end(), cp.jcol.begin(), cp.jcol.end());
            a.tmp.insert(a.tmp.end(), cp.tmp.begin(), cp.tmp.end());
            for (const auto &row : cp.Cd) a.Cd.insert(a.Cd.end(), row.begin(), row.end());
            for (const auto &row : cp.Ce) a.Ce.insert(a.Ce.end(), row.begin(), row.end());
        }
        mgr_species_view &w = sv[l];
        w.num = ld.num;
        w.inverse_mass = ld.inverse_mass;
        w.nlev = ld.nlev;
        w.nrad = ld.nrad;
        w.energy = ld.energy.data();
        w.weight = ld.weight.data();
        w.irad = a.irad.data();
        w.jrad = a.jrad.data();
        w.A = ld.A.data();
        w.Bs = ld.Bs.data();
        w.Ba = ld.Ba.data();
        w.frequency = ld.frequency.data();
        w.ncolpar = (int64_t)ld.colpar.size();
        w.colpar_num = a.cp_num.data();
        w.colpar_orth_or_para = a.cp_op.c_str();
        w.colpar_ncol = a.cp_ncol.data();
        w.colpar_ntmp = a.cp_ntmp.data();
        w.icol = a.icol.data();
        w.jcol = a.jcol.data();
        w.tmp = a.tmp.data();
        w.Cd = a.Cd.data();
        w.Ce = a.Ce.data();
        w.roots = ls.quadrature.roots.data();
        w.qweights = ls.quadrature.weights.data();
    }

    mgr_model_view v{};
    v.ncells = (int64_t)N;
    v.nrays = (int64_t)R;
    v.nboundary = (int64_t)b2c.size();
    v.nspecs = (int64_t)nspecs;
    v.nquads = (int64_t)nq;
    v.pop_prec = pop_prec();
    v.max_width_fraction = parameters.max_width_fraction;
    v.n_off_diag = parameters.n_off_diag;
    v.position = pos.data();
    v.velocity = vel.data();
    v.n_neighbors = n_nb64.data();
    v.neighbors = nbs64.data();
    v.rays = rays.data();
    v.weights = geometry.rays.weights.data();
    v.boundary2cell_nr = b2c.data();
    v.temperature = thermodynamics.temperature.gas.data();
    v.vturb2 = thermodynamics.turbulence.vturb2.data();
    v.abundance = abundance.data();
    const mgr_species_view &w0 = sv[0];
    v.num = w0.num;
    v.inverse_mass = w0.inverse_mass;
    v.nlev = w0.nlev;
    v.nrad = w0.nrad;
    v.energy = w0.energy;
    v.weight = w0.weight;
    v.irad = w0.irad;
    v.jrad = w0.jrad;
    v.A = w0.A;
    v.Bs = w0.Bs;
    v.Ba = w0.Ba;
    v.frequency = w0.frequency;
    v.ncolpar = w0.ncolpar;
    v.colpar_num = w0.colpar_num;
    v.colpar_orth_or_para = w0.colpar_orth_or_para;
    v.colpar_ncol = w0.colpar_ncol;
    v.colpar_ntmp = w0.colpar_ntmp;
    v.icol = w0.icol;
    v.jcol = w0.jcol;
    v.tmp = w0.tmp;
    v.Cd = w0.Cd;
    v.Ce = w0.Ce;
    v.roots = w0.roots;
    v.qweights = w0.qweights;
    v.nlspecs = (int64_t)S;
    v.more_species = S > 1 ? sv.data() + 1 : nullptr;
    const int rc = mgr_create(&v, device, &ctx_);
    if (rc != MGR_OK) {
        ctx_ = nullptr;
        throw std::runtime_error(std::string("mgr_create failed: ") + mgr_last_error(nullptr));
    }
    if (use_scattering()) check(mgr_set_scattering(ctx_, 1), "set_scattering");
}

bool Simulation::use_scattering() const
{
    return parameters.use_scattering_.is_set && parameters.use_scattering();
}

// Simulation::compute_spectral_discretisation (sim_spectrum.tpp:13-196)
int Simulation::compute_spectral_discretisation()
{
    ensure_context();
    check(mgr_compute_spectral_discretisation(ctx_), "compute_spectral_discretisation");
    pull_spectral();
    specDiscSetting = LineSet;
    return 0;
}

// Simulation::compute_spectral_discretisation_image (sim_spectrum.tpp:206-311)
int Simulation::compute_spectral_discretisation_image(const double width)
{
    ensure_context();
    check(mgr_compute_spectral_discretisation_image(ctx_, width), "compute_spectral_discretisation_image");
    pull_spectral();
    specDiscSetting = ImageSet;
    return 0;
}

// Frequencies::nu, nr_line and the unsorted-index back-maps of either grid
// (sim_spectrum.tpp:153-184, 270-302).
// Lines of every species, species after species: total and first line of l.
size_t Simulation::nlines_total() const
{
    size_t K = 0;
    for (const LineProducingSpecies &ls : lines.lineProducingSpecies) K += ls.linedata.nrad;
    return K;
}

size_t Simulation::first_line(size_t l) const
{
    size_t K = 0;
    for (size_t q = 0; q < l; q++) K += lines.lineProducingSpecies[q].linedata.nrad;
    return K;
}

void Simulation::pull_spectral()
{
    const size_t N = geometry.cells.position.size();
    const size_t K = nlines_total(), Q = lines.lineProducingSpecies[0].quadrature.roots.size(), F = K * Q;
    Double1 nu(N * F);
    std::vector<int64_t> nr(N * K * Q);
    check(mgr_get_frequencies(ctx_, nu.data(), nr.data()), "get_frequencies");
    Frequencies &fr = radiation.frequencies;
    fr.nu.assign(N, Double1(F));
    for (size_t p = 0; p < N; p++)
        for (size_t f = 0; f < F; f++) fr.nu[p][f] = nu[p * F + f];
    fr.appears_in_line_integral.assign(F, false);
    fr.corresponding_l_for_spec.assign(F, (long)F);
    fr.corresponding_k_for_tran.assign(F, (long)F);
    fr.corresponding_z_for_line.assign(F, (long)F);
    size_t index2 = 0;
    for (size_t l = 0; l < lines.lineProducingSpecies.size(); l++) {
        LineProducingSpecies &ls = lines.lineProducingSpecies[l];
        const size_t Kl = ls.linedata.nrad, g0 = first_line(l);
        ls.nr_line.assign(N, Long2(Kl, Long1(Q)));
        for (size_t p = 0; p < N; p++)
            for (size_t k = 0; k < Kl; k++)
                for (size_t z = 0; z < Q; z++) ls.nr_line[p][k][z] = (long)nr[(p * K + g0 + k) * Q + z];
        for (size_t k = 0; k < Kl; k++)
            for (size_t z = 0; z < Q; z++) {
                fr.appears_in_line_integral[index2] = true;
                fr.corresponding_l_for_spec[index2] = (long)l;
                fr.corresponding_k_for_tran[index2] = (long)k;
                fr.corresponding_z_for_line[index2] = (long)z;
                index2++;
            }
    }
}

// Copies the device's boundary intensities into radiation.I_bdy; the
// reference stores them per reduced ray, all rays hold the same values.
void Simulation::pull_boundary_intensities()
{
    const size_t NB = geometry.boundary.boundary2cell_nr.size();
    const size_t F = radiation.frequencies.nu.empty() ? 0 : radiation.frequencies.nu[0].size();
    Double1 I(NB * F);
    check(mgr_get_boundary_intensities(ctx_, I.data()), "get_boundary_intensities");
    const size_t Rr = geometry.rays.rays.size() / 2;
    radiation.I_bdy.assign(Rr, Double2(NB, Double1(F)));
    for (size_t r = 0; r < Rr; r++)
        for (size_t b = 0; b < NB; b++)
            for (size_t f = 0; f < F; f++) radiation.I_bdy[r][b][f] = I[b * F + f];
}

// Simulation::compute_boundary_intensities (sim_radiation.tpp:6-14): CMB on
// every boundary cell
int Simulation::compute_boundary_intensities()
{
    ensure_context();
    check(mgr_compute_boundary_intensities(ctx_, nullptr), "compute_boundary_intensities");
    pull_boundary_intensities();
    return 0;
}

// sim_radiation.tpp:17-40: Planck at the given temperatures; a size mismatch
// returns -1
int Simulation::compute_boundary_intensities(const Double1 &temperatures)
{
    if (temperatures.size() != geometry.boundary.boundary2cell_nr.size()) {
        std::fprintf(stderr, "ERROR : temperatures.size() != nboundary\n");
        return -1;
    }
    ensure_context();
    check(mgr_compute_boundary_intensities(ctx_, temperatures.data()), "compute_boundary_intensities");
    pull_boundary_intensities();
    return 0;
}

// The device holds the species' populations one after the other, each in
// the species' own layout i + p*nlev.
void Simulation::pull_populations()
{
    const size_t N = geometry.cells.position.size();
    size_t tot = 0;
    for (const LineProducingSpecies &ls : lines.lineProducingSpecies) tot += N * ls.linedata.nlev;
    Double1 a(tot), b(tot), c(tot), d(tot);
    check(mgr_get_populations(ctx_, a.data()), "get_populations");
    check(mgr_get_population_history(ctx_, b.data(), c.data(), d.data()), "get_population_history");
    size_t o = 0;
    for (LineProducingSpecies &ls : lines.lineProducingSpecies) {
        const size_t n = N * ls.linedata.nlev;
        ls.population.assign(a.begin() + o, a.begin() + o + n);
        ls.population_prev1.assign(b.begin() + o, b.begin() + o + n);
        ls.population_prev2.assign(c.begin() + o, c.begin() + o + n);
        ls.population_prev3.assign(d.begin() + o, d.begin() + o + n);
        o += n;
    }
    lines.emissivity.resize(N * nlines_total());
    lines.opacity.resize(N * nlines_total());
    check(mgr_get_emissivity_opacity(ctx_, lines.emissivity.data(), lines.opacity.data()),
          "get_emissivity_opacity");
}

void Simulation::push_populations()
{
    Double1 a, b, c, d;
    for (const LineProducingSpecies &ls : lines.lineProducingSpecies) {
        a.insert(a.end(), ls.population.begin(), ls.population.end());
        b.insert(b.end(), ls.population_prev1.begin(), ls.population_prev1.end());
        c.insert(c.end(), ls.population_prev2.begin(), ls.population_prev2.end());
        d.insert(d.end(), ls.population_prev3.begin(), ls.population_prev3.end());
    }
    check(mgr_set_populations(ctx_, a.data()), "set_populations");
    check(mgr_set_population_history(ctx_, b.data(), c.data(), d.data()), "set_population_history");
}

// Simulation::compute_LTE_level_populations (sim_lines.tpp:10-16)
int Simulation::compute_LTE_level_populations()
{
    ensure_context();
    check(mgr_compute_LTE_level_populations(ctx_), "compute_LTE_level_populations");
    const size_t N = geometry.cells.position.size();
    for (LineProducingSpecies &ls : lines.lineProducingSpecies) {
        ls.population_tot.resize(N);
        for (size_t p = 0; p < N; p++) ls.population_tot[p] = chemistry.species.abundance[p][ls.linedata.num];
    }
    pull_populations();
    return 0;
}

void Simulation::pull_radiation()
{
    const size_t N = geometry.cells.position.size();
    const size_t Ktot = nlines_total();
    const size_t F = Ktot * lines.lineProducingSpecies[0].quadrature.roots.size();
    radiation.J.resize(N * F);
    check(mgr_get_J(ctx_, radiation.J.data()), "get_J");
    Double1 Lall(N * Ktot);
    check(mgr_get_lambda(ctx_, Lall.data()), "get_lambda");
    // the element nr = p of every (p, k) of every species, when any was added
    for (size_t l = 0; l < lines.lineProducingSpecies.size(); l++) {
        LineProducingSpecies &ls = lines.lineProducingSpecies[l];
        const size_t K = ls.linedata.nrad, g0 = first_line(l);
        ls.lambda.Ls.assign(N, Double2(K));
        ls.lambda.nr.assign(N, Long2(K));
        ls.lambda.size.assign(N * K, 0);
        for (size_t p = 0; p < N; p++)
            for (size_t k = 0; k < K; k++)
                if (Lall[p * Ktot + g0 + k] != 0.0) {
                    ls.lambda.Ls[p][k] = {Lall[p * Ktot + g0 + k]};
                    ls.lambda.nr[p][k] = {(long)p};
                }
        for (size_t p = 0; p < N; p++)
            for (size_t k = 0; k < K; k++) ls.lambda.size[k + K * p] = ls.lambda.nr[p][k].size();
    }
    LineProducingSpecies &ls = lines.lineProducingSpecies[0];
    const size_t K = ls.linedata.nrad;
    // n_off_diag >= 2: the band elements, merged per emitting cell in the
    // order RayPair::update_Lambda first adds them (ray pairs, then bands m,
    // lower before upper; raypair_lambda.tpp:129-146, lambda.tpp:130-143)
    if (parameters.n_off_diag >= 2) {
        int64_t S = 0;
        check(mgr_lambda_band_slots(ctx_, &S), "lambda_band_slots");
        std::vector<int32_t> cell(N * S);
        Double1 Lb(N * S * K);
        check(mgr_get_lambda_band(ctx_, cell.data(), Lb.data()), "get_lambda_band");
        const int64_t nb = parameters.n_off_diag - 1;
        for (size_t p = 0; p < N; p++)
            for (int64_t r = 0; r < S / (2 * nb); r++)
                for (int64_t m = 0; m < nb; m++)
                    for (int64_t side = 0; side < 2; side++) {
                        const int64_t sl = p * S + (r * 2 + side) * nb + m;
                        if (cell[sl] < 0) continue;
                        for (size_t k = 0; k < K; k++) {
                            Double1 &ls_pk = ls.lambda.Ls[p][k];
                            Long1 &nr_pk = ls.lambda.nr[p][k];
                            if (nr_pk.empty()) {   // the diagonal element comes first
                                ls_pk.push_back(0.0);
                                nr_pk.push_back((long)p);
                            }
                            size_t e = 0;
                            while (e < nr_pk.size() && nr_pk[e] != cell[sl]) e++;
                            if (e == nr_pk.size()) {
                                ls_pk.push_back(0.0);
                                nr_pk.push_back(cell[sl]);
                            }
                            ls_pk[e] += Lb[sl * K + k];
                        }
                    }
    }
    for (size_t p = 0; p < N; p++)
        for (size_t k = 0; k < K; k++) ls.lambda.size[k + K * p] = ls.lambda.nr[p][k].size();
    // Radiation::u with scattering (radiation.cpp:117-126), u[R][f + p*F]
    if (use_scattering()) {
        const size_t Rr = geometry.rays.rays.size() / 2;
        Double1 u(Rr * N * F);
        check(mgr_get_u(ctx_, u.data()), "get_u");
        radiation.u.assign(Rr, Double1(N * F));
        for (size_t r = 0; r < Rr; r++)
            std::copy(u.begin() + r * N * F, u.begin() + (r + 1) * N * F, radiation.u[r].begin());
    }
}

void Simulation::pull_jeff()
{
    const size_t N = geometry.cells.position.size();
    const size_t Ktot = nlines_total();
    Double1 a(N * Ktot), b(N * Ktot), c(N * Ktot);
    check(mgr_get_Jeff(ctx_, a.data(), b.data(), c.data()), "get_Jeff");
    for (size_t l = 0; l < lines.lineProducingSpecies.size(); l++) {
        LineProducingSpecies &ls = lines.lineProducingSpecies[l];
        const size_t K = ls.linedata.nrad, g0 = first_line(l);
        ls.Jlin.assign(N, Double1(K));
        ls.Jeff.assign(N, Double1(K));
        ls.Jdif.assign(N, Double1(K));
        for (size_t p = 0; p < N; p++)
            for (size_t k = 0; k < K; k++) {
                ls.Jlin[p][k] = a[p * Ktot + g0 + k];
                ls.Jeff[p][k] = b[p * Ktot + g0 + k];
                ls.Jdif[p][k] = c[p * Ktot + g0 + k];
            }
    }
}

// LineProducingSpecies::check_for_convergence's figures of every species
void Simulation::pull_convergence()
{
    for (size_t l = 0; l < lines.lineProducingSpecies.size(); l++) {
        mgr_conv c{};
        check(mgr_species_convergence(ctx_, (int64_t)l, &c), "species_convergence");
        LineProducingSpecies &ls = lines.lineProducingSpecies[l];
        ls.fraction_not_converged = c.fraction_not_converged;
        ls.relative_change_mean = c.relative_change_mean;
        ls.relative_change_max = c.relative_change_max;
    }
}

int Simulation::run_radiation_field(int prec)
{
    ensure_context();
    push_populations();
    check(mgr_compute_radiation_field(ctx_, prec), "compute_radiation_field");
    double ms = 0.0;
    mgr_last_radiation_field_ms(ctx_, &ms);
    int64_t sum_ndep = 0, pairs = 0;
    mgr_trace_stats(ctx_, &sum_ndep, &pairs);
    timings.push_back(ms);
    nrpairs.push_back((long)pairs);
    depths.push_back((long)sum_ndep);
    pull_radiation();
    return 0;
}

// Simulation::compute_radiation_field (sim_radiation.tpp:56-338)
int Simulation::compute_radiation_field() { return run_radiation_field(precision); }
// the reference's FP64 CPU path: here the FP64 variant of the GPU solve
int Simulation::compute_radiation_field_cpu() { return run_radiation_field(MGR_FP64); }
int Simulation::compute_radiation_field_gpu() { return run_radiation_field(MGR_FP32); }
// The block and grid arguments of the reference's launch are not needed: the
// kernel maps one workgroup per origin cell.
int Simulation::gpu_compute_radiation_field(size_t, size_t, size_t, double) { return run_radiation_field(MGR_FP32); }
int Simulation::cpu_compute_radiation_field(size_t, size_t, size_t, double) { return run_radiation_field(MGR_FP64); }

int Simulation::gpu_get_device_properties()
{
    int n = 0;
    mgr_device_count(&n);
    std::printf("HIP devices: %d (gfx950 library)\n", n);
    return 0;
}

// Simulation::compute_Jeff (sim_lines.tpp:176-220)
void Simulation::compute_Jeff()
{
    ensure_context();
    push_populations();
    check(mgr_compute_Jeff(ctx_), "compute_Jeff");
    pull_jeff();
}

// Simulation::compute_level_populations_from_stateq (sim_lines.tpp:24-30)
void Simulation::compute_level_populations_from_stateq()
{
    ensure_context();
    mgr_conv conv{};
    check(mgr_compute_level_populations_from_stateq(ctx_, &conv), "compute_level_populations_from_stateq");
    pull_convergence();
    pull_populations();
}

long Simulation::compute_level_populations(const Io &io)
{
    return compute_level_populations(io, true, 1000);   // sim_lines.tpp:42-48
}

// Simulation::compute_level_populations (sim_lines.tpp:61-164): the same
// iteration count, Ng schedule, convergence test and per-iteration writes;
// each iteration is one hipGraph replay on the device.
long Simulation::compute_level_populations(const Io &io, bool use_Ng_acceleration, long max_niterations)
{
    if (specDiscSetting != LineSet) {
        std::fprintf(stderr, "Error: Spectral discretisation was not set for Lines!\n");
        return -1;
    }
    ensure_context();
    const size_t S = lines.lineProducingSpecies.size();
    for (size_t l = 0; l < S; l++) lines.lineProducingSpecies[l].write_populations(io, (long)l, "_rank_0_iteration_0");
    push_populations();
    long iteration = 0, iteration_normal = 0;
    error_mean.clear();
    error_max.clear();
    bool some_not_converged = true;
    while (some_not_converged && iteration < max_niterations) {
        iteration++;
        some_not_converged = false;
        const bool ng = use_Ng_acceleration && (iteration_normal == 4);
        mgr_conv conv{};
        check(mgr_ali_iteration(ctx_, ng ? 1 : 0, precision, &conv), "ALI iteration");
        iteration_normal = ng ? 0 : iteration_normal + 1;
        if (S > 1) pull_convergence();
        for (size_t l = 0; l < S; l++) {   // sim_lines.tpp:130-150, per species
            LineProducingSpecies &ls = lines.lineProducingSpecies[l];
            if (S == 1) {
                ls.fraction_not_converged = conv.fraction_not_converged;
                ls.relative_change_mean = conv.relative_change_mean;
                ls.relative_change_max = conv.relative_change_max;
            }
            error_mean.push_back(ls.relative_change_mean);
            error_max.push_back(ls.relative_change_max);
            if (convergence_criterion == MGR_CRIT_MAX) {
                if (ls.relative_change_max >= pop_prec()) some_not_converged = true;
            } else if (ls.fraction_not_converged > 0.005) {
                some_not_converged = true;
            }
        }
        pull_populations();
        if (!ng) pull_jeff();
        for (size_t l = 0; l < S; l++)
            lines.lineProducingSpecies[l].write_populations(io, (long)l, "_rank_0_iteration_" + std::to_string(iteration));
    }
    pull_radiation();
    return iteration;
}

// Simulation::compute_and_write_image (sim_radiation.tpp:345-478): the
// intensities leaving every cell's rest-frame ray pair along ray r, from the
// Model's current populations; a ray outside [0, nrays/2) leaves them zero
// (the reference's "This rank is not computing an image.").
int Simulation::compute_and_write_image(const Io &io, const long r)
{
    if (specDiscSetting != ImageSet) {
        std::fprintf(stderr, "Error: Spectral discretisation was not set for Images!\n");
        return -1;
    }
    ensure_context();
    Image image(r, parameters);
    const size_t N = geometry.cells.position.size();
    const size_t F = radiation.frequencies.nu.empty() ? 0 : radiation.frequencies.nu[0].size();
    const long Rr = (long)geometry.rays.rays.size() / 2;
    if (r >= 0 && r < Rr) {
        push_populations();
        Double1 Im(N * F), Ip(N * F);
        check(mgr_compute_image(ctx_, r, image_precision, Im.data(), Ip.data()), "compute_image");
        for (size_t p = 0; p < N; p++)
            for (size_t f = 0; f < F; f++) {
                image.I_m[p][f] = Im[p * F + f];
                image.I_p[p][f] = Ip[p * F + f];
            }
    } else {
        std::fprintf(stderr, "This rank is not computing an image.\n");
    }
    image.set_coordinates(geometry);
    image.write(io);
    return 0;
}

Long2 Simulation::get_npoints_on_rays_rest()
{
    ensure_context();
    const size_t N = geometry.cells.position.size(), Rr = geometry.rays.rays.size() / 2;
    std::vector<int64_t> np(N * Rr);
    check(mgr_get_npoints_on_rays_rest(ctx_, np.data()), "get_npoints_on_rays_rest");
    Long2 out(Rr, Long1(N));
    for (size_t r = 0; r < Rr; r++)
        for (size_t o = 0; o < N; o++) out[r][o] = (long)np[r * N + o];
    return out;
}

Long2 Simulation::get_npoints_on_rays_comoving()
{
    ensure_context();
    check(mgr_trace(ctx_), "trace");
    const size_t N = geometry.cells.position.size(), Rr = geometry.rays.rays.size() / 2;
    std::vector<int64_t> np(N * Rr);
    check(mgr_get_npoints_on_rays(ctx_, np.data()), "get_npoints_on_rays");
    Long2 out(Rr, Long1(N));
    for (size_t r = 0; r < Rr; r++)
        for (size_t o = 0; o < N; o++) out[r][o] = (long)np[r * N + o];
    return out;
}

}  // namespace magritte
