import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "pdynamo-mirror_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; parity tests of the HIP path")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


class DHFRFixture:
    """The 23,558-atom DHFR non-bonding inputs (tests/golden/dhfr_nb.npz, written by make_dhfr.py from the
    reference's benchmark data) and the reference's published first-call values."""

    def __init__(self, path):
        from pdynamo_amd.synthetic import MMSystem
        with np_load(path) as z:
            self.mm = MMSystem("dhfr", z["coordinates3"], z["charges"], z["ljtypes"], int(z["ntypes"]), z["tableA"],
                               z["tableB"], z["exclusions"], z["interactions14"])
            self.tableA14, self.tableB14 = z["tableA14"], z["tableB14"]
            self.scale14 = float(z["scale14"])
            self.published_terms = dict(zip(("MM/MM Elect.", "MM/MM LJ", "MM/MM 1-4 Elect.", "MM/MM 1-4 LJ"),
                                            (float(v) for v in z["published_terms"])))
            self.published_pairs, self.published_pairs14 = (int(v) for v in z["published_pairs"])
            self.published_image_terms = dict(zip(("MM/MM Image Elect.", "MM/MM Image LJ"),
                                                  (float(v) for v in z["published_image_terms"])))
            self.published_images, self.published_image_pairs = (int(v) for v in z["published_images"])
            self.cell = tuple(float(v) for v in z["cell"])
            self.transformations = (np.eye(3)[None], np.zeros((1, 3)))

    def system(self, periodic=False, **model_options):
        from pdynamo_amd import MMAtomContainer, NBModelABFS, System
        from pdynamo_amd.Parameters import LJParameterContainer_FromTableCoefficients
        m = self.mm
        lj = LJParameterContainer_FromTableCoefficients(m.ntypes, m.tableA, m.tableB)
        lj14 = LJParameterContainer_FromTableCoefficients(m.ntypes, self.tableA14, self.tableB14)
        s = System.FromMMParameters(m.coordinates3, MMAtomContainer(m.charges, m.ljtypes), lj,
                                    exclusions=m.exclusions, interactions14=m.interactions14, ljParameters14=lj14,
                                    electrostaticScale14=self.scale14, label="DHFR")
        if periodic:
            from pdynamo_amd.Symmetry import CrystalClassCubic
            s.DefineSymmetry(crystalClass=CrystalClassCubic(), a=self.cell[0])
        s.DefineNBModel(NBModelABFS(**model_options))
        return s


def np_load(path):
    import numpy as np
    return np.load(path, allow_pickle=False)


@pytest.fixture(scope="session")
def dhfr():
    return DHFRFixture(os.path.join(ROOT, "tests", "golden", "dhfr_nb.npz"))
