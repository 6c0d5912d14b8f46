"""Write test fixtures from the reference's own scenario decks.

Reads the input decks under /root/reference/tests/{hydrostaticPressure,dambreak,waves}
(data files only, nothing of the reference is run) and stores them as compressed
arrays next to this script, so tests on machines without the reference use the
same inputs:

  hydrostatic_init.npz : FluidParticleInit.dat (relaxed, non-lattice fluid) +
                         FixedBoundaryParticle.dat + parameters.dat values
  dambreak.npz         : FluidParticle.dat + FixedBoundaryParticle.dat +
                         parameters.dat (nearNeighbor = 91: the K = 40 cap binds)
  waves.npz            : FluidParticle.dat + Fixed/MovingBoundaryParticle.dat +
                         parameters.dat (piston wavemaker)
  hydrostatic_probe.npz: recorded time series PointPressure.dat / TotalKineticEnergy.dat
  hydrostatic_start.npz: FluidParticle.dat (the lattice the recorded run started from) +
                         FixedBoundaryParticle.dat + parameters.dat values; the recorded
                         run's output() appended its state at step
                         NumberOfInitializationSteps = 30000 as FluidParticleInit.dat
                         (hydrostatic_init.npz)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from uph_amd.scenes import read_deck, read_parameters  # noqa: E402

REF = "/root/reference/tests"


def save(name, d, fluid_file="FluidParticle.dat"):
    src = os.path.join(REF, d)
    prm = read_parameters(os.path.join(src, "parameters.dat"))
    fl = read_deck(os.path.join(src, fluid_file))
    fx = read_deck(os.path.join(src, "FixedBoundaryParticle.dat"))
    mv = read_deck(os.path.join(src, "MovingBoundaryParticle.dat"))
    keys = np.array(list(prm.keys()))
    vals = np.array(list(prm.values()))
    np.savez_compressed(os.path.join(HERE, name), fluid=fl, fixed=fx, moving=mv,
                        param_names=keys, param_values=vals)
    print(name, fl.shape, fx.shape, mv.shape)


def main():
    save("hydrostatic_init.npz", "hydrostaticPressure", "FluidParticleInit.dat")
    save("hydrostatic_start.npz", "hydrostaticPressure")
    save("dambreak.npz", "dambreak")
    save("waves.npz", "waves")
    pp = np.loadtxt(os.path.join(REF, "hydrostaticPressure", "PointPressure.dat"))
    ke = np.loadtxt(os.path.join(REF, "hydrostaticPressure", "TotalKineticEnergy.dat"))
    np.savez_compressed(os.path.join(HERE, "hydrostatic_probe.npz"), point_pressure=pp,
                        kinetic_energy=ke)


if __name__ == "__main__":
    main()
