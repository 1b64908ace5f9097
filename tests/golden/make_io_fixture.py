"""Writes tests/golden/gro_excerpt.gro: whole molecules cut from the reference's input file
input_files/glycine_1500_solv_min.gro (data, copied line for line).

The excerpt keeps glycines 999-1002 (atoms 9981-10020, where the atom number runs into the
atom-name field from 10000 on), waters 1501-1510 (atoms 15001-15030) and the last water
(atoms 45253-45255), the original atom numbers and the box line.

    python tests/golden/make_io_fixture.py
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/input_files/glycine_1500_solv_min.gro"
RANGES = [(9981, 10020), (15001, 15030), (45253, 45255)]


def main():
    lines = open(SRC).read().split("\n")
    atoms = [lines[1 + k] for a, b in RANGES for k in range(a, b + 1)]
    box = lines[2 + int(lines[1])]
    with open(os.path.join(HERE, "gro_excerpt.gro"), "w") as f:
        f.write(lines[0] + "\n")
        f.write(f"{len(atoms):5d}\n")
        for l in atoms:
            f.write(l + "\n")
        f.write(box + "\n")
    print(f"wrote {len(atoms)} atoms")


if __name__ == "__main__":
    main()
