"""CPU side of the solution files: the tables carry the species names in the
mechanism's SPECIES order (they name the scalars in the file), and the
library exports the writer and the reader."""
from corefl_amd import _abi, load_air5_tables, load_tables


def names(t):
    return [t.species_name[i].value.decode() for i in range(t.n_spec)]


def test_species_names_follow_the_species_section():
    # data/h2_li2004/mechanism.inp and data/air5_park/mechanism.inp list these, in this order
    assert names(load_tables()) == ["H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"]
    assert names(load_air5_tables()) == ["N2", "O2", "NO", "N", "O"]


def test_solution_file_entry_points_are_exported():
    lib = _abi.load()
    for sym in ("corefl_plt_write", "corefl_plt_read"):
        assert hasattr(lib, sym)
