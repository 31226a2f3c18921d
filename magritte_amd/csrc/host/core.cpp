// core.cpp -- the Python module `core`, with the class and member names of the
// reference's bindings (src/bindings/pyBindings.cpp:25-511) for the parts of
// the model that the radiation-field / level-population path reads or writes.
// Vectors convert to and from Python lists; the two vectors of classes are
// bound opaque so that items are edited in place, as in the reference
// (pyBindings.cpp:21-34).
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>
#include <pybind11/stl_bind.h>

#include "simulation.hpp"
#ifdef MGR_WITH_HDF5
#include "io_hdf5.hpp"
#endif

namespace py = pybind11;
using namespace magritte;

PYBIND11_MAKE_OPAQUE(std::vector<LineProducingSpecies>);
PYBIND11_MAKE_OPAQUE(std::vector<CollisionPartner>);

PYBIND11_MODULE(core, module)
{
    module.doc() = "gfx950 drop-in of the Magritte core module (radiation field and level populations)";

    py::bind_vector<std::vector<LineProducingSpecies>>(module, "vLineProducingSpecies");
    py::bind_vector<std::vector<CollisionPartner>>(module, "vCollisionPartner");

    py::class_<Io>(module, "Io");
    py::class_<IoText, Io>(module, "IoText")
        .def_readonly("io_file", &IoText::io_file)
        .def(py::init<const std::string &>());
#ifdef MGR_WITH_HDF5
    // IoPython("hdf5", file) (pyBindings.cpp:46-56), native HDF5
    py::class_<IoHDF5, Io>(module, "IoHDF5")
        .def_readonly("io_file", &IoHDF5::io_file)
        .def(py::init<const std::string &>());
    py::class_<IoPython, IoHDF5>(module, "IoPython")
        .def_readonly("implementation", &IoPython::implementation)
        .def_readonly("io_file", &IoPython::io_file)
        .def(py::init<const std::string &, const std::string &>())
        .def("read_number", [](const IoPython &io, const std::string &name) {
            size_t v = 0;
            io.read_number(name, v);
            return v;
        });
#endif

    py::class_<Model>(module, "Model")
        .def_readwrite("parameters", &Model::parameters)
        .def_readwrite("geometry", &Model::geometry)
        .def_readwrite("chemistry", &Model::chemistry)
        .def_readwrite("lines", &Model::lines)
        .def_readwrite("thermodynamics", &Model::thermodynamics)
        .def_readwrite("radiation", &Model::radiation)
        .def(py::init())
        .def("read", &Model::read)
        .def("write", &Model::write);

    py::class_<Parameters>(module, "Parameters")
        .def(py::init())
        .def_readwrite("n_off_diag", &Parameters::n_off_diag)
        .def_readwrite("max_width_fraction", &Parameters::max_width_fraction)
        .def("set_ncells", &Parameters::set_ncells)
        .def("set_nrays", &Parameters::set_nrays)
        .def("set_nrays_red", &Parameters::set_nrays_red)
        .def("set_order_min", &Parameters::set_order_min)
        .def("set_order_max", &Parameters::set_order_max)
        .def("set_nboundary", &Parameters::set_nboundary)
        .def("set_nfreqs", &Parameters::set_nfreqs)
        .def("set_nfreqs_red", &Parameters::set_nfreqs_red)
        .def("set_nspecs", &Parameters::set_nspecs)
        .def("set_nlspecs", &Parameters::set_nlspecs)
        .def("set_nlines", &Parameters::set_nlines)
        .def("set_nquads", &Parameters::set_nquads)
        .def("set_pop_prec", &Parameters::set_pop_prec)
        .def("set_use_scattering", &Parameters::set_use_scattering)
        .def("set_spherical_symmetry", &Parameters::set_spherical_symmetry)
        .def("set_adaptive_ray_tracing", &Parameters::set_adaptive_ray_tracing)
        .def("ncells", &Parameters::ncells)
        .def("nrays", &Parameters::nrays)
        .def("nrays_red", &Parameters::nrays_red)
        .def("order_min", &Parameters::order_min)
        .def("order_max", &Parameters::order_max)
        .def("nboundary", &Parameters::nboundary)
        .def("nfreqs", &Parameters::nfreqs)
        .def("nfreqs_red", &Parameters::nfreqs_red)
        .def("nspecs", &Parameters::nspecs)
        .def("nlspecs", &Parameters::nlspecs)
        .def("nlines", &Parameters::nlines)
        .def("nquads", &Parameters::nquads)
        .def("pop_prec", &Parameters::pop_prec)
        .def("use_scattering", &Parameters::use_scattering)
        .def("spherical_symmetry", &Parameters::spherical_symmetry)
        .def("adaptive_ray_tracing", &Parameters::adaptive_ray_tracing)
        .def("read", &Parameters::read)
        .def("write", &Parameters::write);

    py::class_<Geometry>(module, "Geometry")
        .def_readwrite("cells", &Geometry::cells)
        .def_readwrite("rays", &Geometry::rays)
        .def_readwrite("boundary", &Geometry::boundary)
        .def(py::init())
        .def("read", &Geometry::read)
        .def("write", &Geometry::write);

    py::class_<Cells>(module, "Cells")
        .def_readwrite("position", &Cells::position)
        .def_readwrite("velocity", &Cells::velocity)
        .def_readwrite("n_neighbors", &Cells::n_neighbors)
        .def_readwrite("neighbors", &Cells::neighbors)
        .def(py::init())
        .def("read", &Cells::read)
        .def("write", &Cells::write);

    py::class_<Rays>(module, "Rays")
        .def_readwrite("rays", &Rays::rays)
        .def_readwrite("weights", &Rays::weights)
        .def_readwrite("antipod", &Rays::antipod)
        .def_readwrite("nrays", &Rays::nrays)
        .def_readwrite("half_nrays", &Rays::half_nrays)
        .def(py::init())
        .def("read", &Rays::read)
        .def("write", &Rays::write);

    py::enum_<BoundaryCondition>(module, "BoundaryCondition")
        .value("Zero", Zero)
        .value("Thermal", Thermal)
        .value("CMB", CMB)
        .export_values();

    py::class_<Boundary>(module, "Boundary")
        .def_readwrite("boundary2cell_nr", &Boundary::boundary2cell_nr)
        .def_readwrite("cell2boundary_nr", &Boundary::cell2boundary_nr)
        .def_readwrite("boundary", &Boundary::boundary)
        .def_readwrite("boundary_condition", &Boundary::boundary_condition)
        .def_readwrite("boundary_temperature", &Boundary::boundary_temperature)
        .def(py::init())
        .def("read", &Boundary::read)
        .def("write", &Boundary::write);

    py::class_<Thermodynamics>(module, "Thermodynamics")
        .def_readwrite("temperature", &Thermodynamics::temperature)
        .def_readwrite("turbulence", &Thermodynamics::turbulence)
        .def(py::init())
        .def("read", &Thermodynamics::read)
        .def("write", &Thermodynamics::write);

    py::class_<Temperature>(module, "Temperature")
        .def_readwrite("gas", &Temperature::gas)
        .def(py::init());

    py::class_<Turbulence>(module, "Turbulence")
        .def_readwrite("vturb2", &Turbulence::vturb2)
        .def(py::init());

    py::class_<Chemistry>(module, "Chemistry")
        .def_readwrite("species", &Chemistry::species)
        .def(py::init())
        .def("read", &Chemistry::read)
        .def("write", &Chemistry::write);

    py::class_<Species>(module, "Species")
        .def_readwrite("abundance", &Species::abundance)
        .def_readwrite("sym", &Species::sym)
        .def(py::init())
        .def("read", &Species::read)
        .def("write", &Species::write);

    py::class_<Lines>(module, "Lines")
        .def_readwrite("lineProducingSpecies", &Lines::lineProducingSpecies)
        .def_readwrite("emissivity", &Lines::emissivity)
        .def_readwrite("opacity", &Lines::opacity)
        .def_readwrite("line", &Lines::line)
        .def_readwrite("line_index", &Lines::line_index)
        .def(py::init<>())
        .def("read", &Lines::read)
        .def("write", &Lines::write);

    py::class_<LineProducingSpecies>(module, "LineProducingSpecies")
        .def_readwrite("linedata", &LineProducingSpecies::linedata)
        .def_readwrite("quadrature", &LineProducingSpecies::quadrature)
        .def_readwrite("Lambda", &LineProducingSpecies::lambda)   // "lambda" is a Python keyword
        .def_readwrite("Jeff", &LineProducingSpecies::Jeff)
        .def_readwrite("Jlin", &LineProducingSpecies::Jlin)
        .def_readwrite("Jdif", &LineProducingSpecies::Jdif)
        .def_readwrite("nr_line", &LineProducingSpecies::nr_line)
        .def_readwrite("population", &LineProducingSpecies::population)
        .def_readwrite("population_tot", &LineProducingSpecies::population_tot)
        .def_readwrite("population_prev1", &LineProducingSpecies::population_prev1)
        .def_readwrite("population_prev2", &LineProducingSpecies::population_prev2)
        .def_readwrite("population_prev3", &LineProducingSpecies::population_prev3)
        .def_readwrite("ncells", &LineProducingSpecies::ncells)
        .def_readwrite("relative_change_mean", &LineProducingSpecies::relative_change_mean)
        .def_readwrite("relative_change_max", &LineProducingSpecies::relative_change_max)
        .def_readwrite("fraction_not_converged", &LineProducingSpecies::fraction_not_converged)
        .def(py::init<>())
        .def("read", &LineProducingSpecies::read)
        .def("write", &LineProducingSpecies::write)
        .def("read_populations", &LineProducingSpecies::read_populations)
        .def("write_populations", &LineProducingSpecies::write_populations)
        .def("index", &LineProducingSpecies::index);

    py::class_<Lambda>(module, "Lambda")
        .def_readwrite("Ls", &Lambda::Ls)
        .def_readwrite("nr", &Lambda::nr)
        .def_readwrite("size", &Lambda::size)
        .def_readwrite("Lss", &Lambda::Lss)
        .def_readwrite("nrs", &Lambda::nrs)
        .def(py::init<>());

    py::class_<Quadrature>(module, "Quadrature")
        .def_readwrite("roots", &Quadrature::roots)
        .def_readwrite("weights", &Quadrature::weights)
        .def(py::init<>())
        .def("read", &Quadrature::read)
        .def("write", &Quadrature::write);

    py::class_<Linedata>(module, "Linedata")
        .def_readwrite("num", &Linedata::num)
        .def_readwrite("sym", &Linedata::sym)
        .def_readwrite("inverse_mass", &Linedata::inverse_mass)
        .def_readwrite("nlev", &Linedata::nlev)
        .def_readwrite("nrad", &Linedata::nrad)
        .def_readwrite("irad", &Linedata::irad)
        .def_readwrite("jrad", &Linedata::jrad)
        .def_readwrite("energy", &Linedata::energy)
        .def_readwrite("weight", &Linedata::weight)
        .def_readwrite("frequency", &Linedata::frequency)
        .def_readwrite("A", &Linedata::A)
        .def_readwrite("Ba", &Linedata::Ba)
        .def_readwrite("Bs", &Linedata::Bs)
        .def_readwrite("ncolpar", &Linedata::ncolpar)
        .def_readwrite("colpar", &Linedata::colpar)
        .def(py::init<>())
        .def("read", &Linedata::read)
        .def("write", &Linedata::write);

    py::class_<CollisionPartner>(module, "CollisionPartner")
        .def_readwrite("num_col_partner", &CollisionPartner::num_col_partner)
        .def_readwrite("orth_or_para_H2", &CollisionPartner::orth_or_para_H2)
        .def_readwrite("ntmp", &CollisionPartner::ntmp)
        .def_readwrite("ncol", &CollisionPartner::ncol)
        .def_readwrite("icol", &CollisionPartner::icol)
        .def_readwrite("jcol", &CollisionPartner::jcol)
        .def_readwrite("tmp", &CollisionPartner::tmp)
        .def_readwrite("Ce", &CollisionPartner::Ce)
        .def_readwrite("Cd", &CollisionPartner::Cd)
        .def(py::init<>())
        .def("read", &CollisionPartner::read)
        .def("write", &CollisionPartner::write);

    py::class_<Radiation>(module, "Radiation")
        .def_readwrite("frequencies", &Radiation::frequencies)
        .def_readwrite("u", &Radiation::u)
        .def_readwrite("v", &Radiation::v)
        .def_readwrite("I_bdy", &Radiation::I_bdy)
        .def_readwrite("J", &Radiation::J)
        .def(py::init())
        .def("read", &Radiation::read)
        .def("write", &Radiation::write)
        .def("get_J", &Radiation::get_J);

    py::class_<Frequencies>(module, "Frequencies")
        .def_readwrite("nu", &Frequencies::nu)
        .def_readwrite("appears_in_line_integral", &Frequencies::appears_in_line_integral)
        .def_readwrite("corresponding_l_for_spec", &Frequencies::corresponding_l_for_spec)
        .def_readwrite("corresponding_k_for_tran", &Frequencies::corresponding_k_for_tran)
        .def_readwrite("corresponding_z_for_line", &Frequencies::corresponding_z_for_line)
        .def(py::init())
        .def("write", &Frequencies::write);

    // Image (pyBindings.cpp:497-507)
    py::class_<Image>(module, "Image")
        .def_readonly("ImX", &Image::ImX)
        .def_readonly("ImY", &Image::ImY)
        .def_readonly("I_p", &Image::I_p)
        .def_readonly("I_m", &Image::I_m)
        .def(py::init<const long, const Parameters &>())
        .def("write", &Image::write)
        .def("set_coordinates", &Image::set_coordinates);

    py::class_<Simulation, Model>(module, "Simulation")
        .def(py::init<>())
        .def_readonly("error_max", &Simulation::error_max)
        .def_readonly("error_mean", &Simulation::error_mean)
        .def_readonly("timings", &Simulation::timings)
        .def_readonly("nrpairs", &Simulation::nrpairs)
        .def_readonly("depths", &Simulation::depths)
        .def_readwrite("device", &Simulation::device)
        .def_readwrite("precision", &Simulation::precision)
        .def_readwrite("convergence_criterion", &Simulation::convergence_criterion)
        .def("gpu_get_device_properties", &Simulation::gpu_get_device_properties)
        .def("gpu_compute_radiation_field", &Simulation::gpu_compute_radiation_field)
        .def("compute_radiation_field_gpu", &Simulation::compute_radiation_field_gpu)
        .def("cpu_compute_radiation_field", &Simulation::cpu_compute_radiation_field)
        .def("compute_radiation_field_cpu", &Simulation::compute_radiation_field_cpu)
        .def("compute_Jeff", &Simulation::compute_Jeff)
        .def("compute_level_populations_from_stateq", &Simulation::compute_level_populations_from_stateq)
        .def("compute_spectral_discretisation", &Simulation::compute_spectral_discretisation)
        .def("compute_spectral_discretisation_image", &Simulation::compute_spectral_discretisation_image)
        .def("compute_and_write_image", &Simulation::compute_and_write_image)
        .def_readwrite("image_precision", &Simulation::image_precision)
        .def("compute_boundary_intensities",
             (int (Simulation::*)(const Double1 &)) & Simulation::compute_boundary_intensities)
        .def("compute_boundary_intensities", (int (Simulation::*)()) & Simulation::compute_boundary_intensities)
        .def("compute_LTE_level_populations", &Simulation::compute_LTE_level_populations)
        .def("compute_radiation_field", &Simulation::compute_radiation_field)
        .def("compute_level_populations", (long (Simulation::*)(const Io &)) & Simulation::compute_level_populations)
        .def("compute_level_populations",
             (long (Simulation::*)(const Io &, bool, long)) & Simulation::compute_level_populations)
        .def("get_npoints_on_rays_comoving", &Simulation::get_npoints_on_rays_comoving)
        .def("get_npoints_on_rays_rest", &Simulation::get_npoints_on_rays_rest)
        .def("reset_device", &Simulation::reset_device);

    module.attr("FP32") = MGR_FP32;
    module.attr("FP64") = MGR_FP64;
    module.attr("CRIT_FRACTION") = MGR_CRIT_FRACTION;
    module.attr("CRIT_MAX") = MGR_CRIT_MAX;
}
