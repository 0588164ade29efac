"""Per-element transport constants (Z, standard atomic weight A, mean excitation J).

Values restate the element table of PyRITE
(``src/pyrite/materials/_transport_data.py:11-55``): CIAAW 2024 standard atomic
weights and PDG mean excitation energies.  They are physical constants, kept
identical because the stopping-power crossover and every transport table are
functions of them.
"""

TRANSPORT_ELEMENTS = {
    "H": {"Z": 1, "A": 1.008, "J_keV": 0.0218},
    "Li": {"Z": 3, "A": 6.94, "J_keV": 0.040},
    "B": {"Z": 5, "A": 10.81, "J_keV": 0.076},
    "C": {"Z": 6, "A": 12.011, "J_keV": 0.078},
    "N": {"Z": 7, "A": 14.007, "J_keV": 0.082},
    "O": {"Z": 8, "A": 15.999, "J_keV": 0.095},
    "F": {"Z": 9, "A": 18.998403163, "J_keV": 0.115},
    "Na": {"Z": 11, "A": 22.98976928, "J_keV": 0.149},
    "Al": {"Z": 13, "A": 26.982, "J_keV": 0.166},
    "Si": {"Z": 14, "A": 28.085, "J_keV": 0.173},
    "P": {"Z": 15, "A": 30.974, "J_keV": 0.173},
    "S": {"Z": 16, "A": 32.06, "J_keV": 0.180},
    "Ar": {"Z": 18, "A": 39.948, "J_keV": 0.188},
    "Ca": {"Z": 20, "A": 40.078, "J_keV": 0.191},
    "Ti": {"Z": 22, "A": 47.867, "J_keV": 0.233},
    "V": {"Z": 23, "A": 50.9415, "J_keV": 0.245},
    "Cr": {"Z": 24, "A": 51.9961, "J_keV": 0.257},
    "Mn": {"Z": 25, "A": 54.938044, "J_keV": 0.272},
    "Fe": {"Z": 26, "A": 55.845, "J_keV": 0.286},
    "Ni": {"Z": 28, "A": 58.6934, "J_keV": 0.311},
    "Cu": {"Z": 29, "A": 63.546, "J_keV": 0.322},
    "Ga": {"Z": 31, "A": 69.723, "J_keV": 0.334},
    "Ge": {"Z": 32, "A": 72.630, "J_keV": 0.350},
    "As": {"Z": 33, "A": 74.921595, "J_keV": 0.347},
    "Se": {"Z": 34, "A": 78.971, "J_keV": 0.348},
    "Zr": {"Z": 40, "A": 91.224, "J_keV": 0.393},
    "Nb": {"Z": 41, "A": 92.906, "J_keV": 0.417},
    "Mo": {"Z": 42, "A": 95.95, "J_keV": 0.424},
    "Pd": {"Z": 46, "A": 106.42, "J_keV": 0.477},
    "Cd": {"Z": 48, "A": 112.414, "J_keV": 0.469},
    "Te": {"Z": 52, "A": 127.60, "J_keV": 0.485},
    "I": {"Z": 53, "A": 126.90447, "J_keV": 0.491},
    "Cs": {"Z": 55, "A": 132.90545196, "J_keV": 0.488},
    "Ba": {"Z": 56, "A": 137.327, "J_keV": 0.491},
    "Hf": {"Z": 72, "A": 178.49, "J_keV": 0.705},
    "Ta": {"Z": 73, "A": 180.94788, "J_keV": 0.718},
    "W": {"Z": 74, "A": 183.84, "J_keV": 0.727},
    "Re": {"Z": 75, "A": 186.207, "J_keV": 0.736},
    "Pt": {"Z": 78, "A": 195.08, "J_keV": 0.790},
    "Au": {"Z": 79, "A": 196.966569, "J_keV": 0.79},
    "Pb": {"Z": 82, "A": 207.2, "J_keV": 0.823},
    "Bi": {"Z": 83, "A": 208.98040, "J_keV": 0.823},
}
