"""CPU tests of the nuclide JSON loader (reference layout, ``//`` comments allowed)."""
import json

import pytest

from pyvoxeldosimetry_amd.dose_kernels import NUCLIDES, load_config, nuclide_from_config, strip_json_comments

# A nuclide file written the way the reference's Lu177/Lu177.json and Ga68/Ga68.json
# are (trailing // comments after values), values as this package states them.
LU177_WITH_COMMENTS = """{
  "metadata": {"created_by": "someone", "url": "http://example.org/a//b"},
  "nuclide": {
    "name": "Lutetium-177",
    "symbol": "Lu177",
    "half_life": 161.52,  // hours
    "particle_energies": {
      "beta-_max": [0.498, 0.385, 0.177],  // MeV
      "beta-_abundance": [0.795, 0.089, 0.116],
      "gamma_lines": [
        {"energy": 0.208, "intensity": 0.111},
        {"energy": 0.113, "intensity": 0.062}
      ]
    }
  },
  "kernel": {
    "voxel_size": 1.0,  // mm
    "grid_size": [81, 81, 81],
    "scaling_factor": 1.0e-12
  }
}"""

GA68_WITH_COMMENTS = """{
  "nuclide": {
    "name": "Gallium-68", "symbol": "Ga68",
    "half_life": 1.128,  // hours
    "particle_energies": {
      "beta+_max": 1.899,  // MeV
      "annihilation": 0.511,  // MeV
      "gamma_lines": [{"energy": 1.077, "intensity": 0.03}]
    }
  },
  "kernel": {"grid_size": [151, 151, 151],  // larger grid due to higher energy
             "scaling_factor": 1.0e-12}
}"""


def test_plain_json_load_rejects_comments_and_the_stripper_fixes_it():
    with pytest.raises(json.JSONDecodeError):
        json.loads(LU177_WITH_COMMENTS)
    cfg = json.loads(strip_json_comments(LU177_WITH_COMMENTS))
    assert cfg["metadata"]["url"] == "http://example.org/a//b"      # '//' inside a string is kept
    assert cfg["nuclide"]["half_life"] == 161.52


def test_escaped_quotes_do_not_end_a_string():
    text = '{"a": "x\\\\" , "b": "say \\"//not a comment\\""} // tail'
    assert json.loads(strip_json_comments(text)) == {"a": "x\\", "b": 'say "//not a comment"'}


@pytest.mark.parametrize("text,symbol", [(LU177_WITH_COMMENTS, "Lu177"), (GA68_WITH_COMMENTS, "Ga68")])
def test_reference_layout_files_give_the_packaged_constants(tmp_path, text, symbol):
    p = tmp_path / f"{symbol}.json"
    p.write_text(text)
    got = nuclide_from_config(load_config(p))
    want = NUCLIDES[symbol]
    for f in ("symbol", "half_life_h", "beta_max_mev", "beta_abundance", "beta_range_coeff", "gammas",
              "bremsstrahlung", "scaling_factor", "default_grid", "positron", "annihilation_mev"):
        assert getattr(got, f) == getattr(want, f), f


def test_generators_take_a_config_path_and_check_the_symbol(tmp_path):
    from pyvoxeldosimetry_amd.dose_kernels import Ga68KernelGenerator, Lu177KernelGenerator
    p = tmp_path / "Ga68.json"
    p.write_text(GA68_WITH_COMMENTS)
    g = Ga68KernelGenerator("water", config_path=p)
    assert g.nuclide.gammas == ((1.077, 0.03),)
    terms, factor = g.terms()
    assert len(terms) == 3 and factor == 1.0
    with pytest.raises(ValueError, match="describes Ga68"):
        Lu177KernelGenerator("water", config_path=p)


def test_factory_knows_ga68_with_its_default_grid():
    from pyvoxeldosimetry_amd import KernelFactory
    assert "Ga68" in KernelFactory.supported()
    assert KernelFactory._default_grid_sizes["Ga68"] == (151, 151, 151)
