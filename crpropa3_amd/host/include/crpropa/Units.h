// SI units and physical constants with the reference's numeric definitions
// (include/crpropa/Units.h:24-116).  The values have to be identical for the
// propagation to agree bit for bit, so each one is defined by the same
// expression.
#ifndef CRPROPA_AMD_UNITS_H
#define CRPROPA_AMD_UNITS_H

#include <cmath>

namespace crpropa {

static const double meter = 1;
static const double second = 1;
static const double kilogram = 1;
static const double ampere = 1;
static const double kelvin = 1;

static const double newton = 1 * kilogram * meter / second / second;
static const double joule = 1 * newton * meter;
static const double tesla = 1 * newton / ampere / meter;
static const double volt = 1 * kilogram * meter * meter / ampere / second / second / second;

static const double eplus = 1.602176487e-19 * ampere * second;
static const double c_light = 2.99792458e8 * meter / second;
static const double c_squared = c_light * c_light;
static const double amu = 1.660538921e-27 * kilogram;
static const double mass_proton = 1.67262158e-27 * kilogram;
static const double mass_electron = 9.10938291e-31 * kilogram;

static const double gauss = 1e-4 * tesla;
static const double microgauss = 1e-6 * gauss;
static const double nanogauss = 1e-9 * gauss;
static const double muG = microgauss;
static const double nG = nanogauss;

static const double electronvolt = eplus * volt;
static const double eV = electronvolt;
static const double keV = 1e3 * electronvolt;
static const double MeV = 1e6 * electronvolt;
static const double GeV = 1e9 * electronvolt;
static const double TeV = 1e12 * electronvolt;
static const double PeV = 1e15 * electronvolt;
static const double EeV = 1e18 * electronvolt;

static const double au = 149597870700 * meter;
static const double parsec = 648000 / M_PI * au;
static const double kiloparsec = 1e3 * parsec;
static const double megaparsec = 1e6 * parsec;
static const double gigaparsec = 1e9 * parsec;
static const double pc = parsec;
static const double kpc = kiloparsec;
static const double Mpc = megaparsec;
static const double Gpc = gigaparsec;

static const double sec = second;
static const double minute = 60 * second;
static const double hour = 3600 * second;
static const double day = 24 * hour;
static const double year = 365.25 * 24 * hour;
static const double kiloyear = 1e3 * year;
static const double Megayear = 1e6 * year;
static const double Gigayear = 1e9 * year;
static const double yr = year;
static const double kyr = kiloyear;
static const double Myr = Megayear;
static const double Gyr = Gigayear;

static const double kilometer = 1000 * meter;
static const double km = kilometer;

} // namespace crpropa

#endif
