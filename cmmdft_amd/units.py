"""Atomic-unit constants (bohr, hartree, electron mass, kelvin).

The reference takes these from the third-party ``molmod`` package
(``molmod.units`` / ``molmod.constants``, imported at cmmdft/functionals.py:11-12,
cmmdft/system.py:12-13, cmmdft/solver.py:9-10, cmmdft/eos.py:12-13).  That
package is not available here, so the values are restated from the CODATA
set molmod uses.  Every quantity in this project (oracle and HIP path alike)
is expressed with these constants, so GPU-versus-oracle parity does not depend
on their last digits; only the absolute mapping pressure -> chemical
potential does.
"""
import math

# --- base conversions (SI -> atomic units) ---
coulomb = 1.0 / 1.602176462e-19
kilogram = 1.0 / 9.10938188e-31
gram = 1.0e-3 * kilogram
unified = 1.66053873e-27 * kilogram
amu = unified
meter = 1.0 / 0.5291772083e-10
centimeter = 1.0e-2 * meter
angstrom = 1.0e-10 * meter
nanometer = 1.0e-9 * meter
avogadro = 6.0221415e23
joule = 1.0 / 4.35974381e-18
calorie = 4.184 * joule
kjmol = 1.0e3 * joule / avogadro
kcalmol = 1.0e3 * calorie / avogadro
newton = joule / meter
pascal = newton / meter ** 2
bar = 1.0e5 * pascal
kelvin = 1.0

# --- physical constants in atomic units ---
boltzmann = 3.1668154051341965e-06   # hartree / kelvin
planck = 2.0 * math.pi               # h = 2*pi*hbar, hbar = 1


def parse_unit(name):
    """Map a unit name as written in LJ parameter files (e.g. 'angstrom',
    'kcalmol', 'kjmol', 'K') to its value in atomic units."""
    table = {
        'angstrom': angstrom, 'A': angstrom, 'nanometer': nanometer, 'nm': nanometer,
        'bohr': 1.0, 'au': 1.0, 'kcalmol': kcalmol, 'kjmol': kjmol,
        'hartree': 1.0, 'meter': meter,
    }
    if name not in table:
        raise ValueError('unknown unit %r' % name)
    return table[name]
