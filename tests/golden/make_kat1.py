"""Extract the KAT-1 golden vectors from the reference's stored SCF result.

Source (data, read as text): tests/reference/runtime/generated/reference-1/
normalized-result.txt and summary.json of chaeeundad/baandi. The run is the
synthetic Si2 SCF (common::potential(4, 1, true, 801), fcc a=10.261212856717933
Bohr, atoms (0,0,0) and (1/4,1/4,1/4), Ecut 5 Ha, 8 bands, LDA-PZ, Gamma).

Writes tests/golden/kat1.json: converged orbitals (137 x 8 complex, the
reference basis order), eigenvalues, occupations, the energy decomposition and
the Ewald breakdown. Run once in a container that has the reference checkout:

    python tests/golden/make_kat1.py /root/reference
"""
import json
import re
import sys
from pathlib import Path


def main(root):
    base = Path(root) / "tests/reference/runtime/generated/reference-1"
    text = (base / "normalized-result.txt").read_text()
    summary = json.loads((base / "summary.json").read_text())
    at = text.find("orbitals: ComplexMatrix")
    head = re.match(r"orbitals: ComplexMatrix \{ rows: (\d+), cols: (\d+), data: \[", text[at:])
    rows, cols = int(head.group(1)), int(head.group(2))
    start = at + head.end()
    end = text.find("]", start)
    pairs = re.findall(r"Complex \{ re: ([^,]+), im: ([^ ]+) \}", text[start:end])
    assert len(pairs) == rows * cols
    energy = re.search(r"energy: EnergyAudit \{ ([^}]*) \}", text).group(1)
    terms = {k: float(v) for k, v in re.findall(r"(\w+): ([-0-9.e]+)", energy)}
    eig = re.search(r"eigenvalues: \[([^\]]*)\]", text).group(1)
    occ = re.search(r"occupations: \[([^\]]*)\]", text).group(1)
    ew = re.search(r"ewald_breakdown: EwaldEnergy \{ ([^}]*) \}", text).group(1)
    ewald = {k: float(v) for k, v in re.findall(r"(\w+): (?:Hartree\()?([-0-9.e]+)", ew)}
    out = {
        "source": "reference-1/normalized-result.txt (chaeeundad/baandi)",
        "cell_a_bohr": 10.261212856717933,
        "ecut_ha": 5.0,
        "rows": rows,
        "cols": cols,
        "orbitals_colmajor_re_im": [[float(a), float(b)] for a, b in pairs],
        "eigenvalues_ha": [float(v) for v in eig.split(",")],
        "occupations": [float(v) for v in occ.split(",")],
        "energy_terms_ha": terms,
        "ewald": ewald,
        "certified_energy_ha": summary["certified_energy_hartree"],
        "density_shape": [15, 15, 15],
    }
    dest = Path(__file__).with_name("kat1.json")
    dest.write_text(json.dumps(out, indent=1))
    print("wrote", dest, rows, cols)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
