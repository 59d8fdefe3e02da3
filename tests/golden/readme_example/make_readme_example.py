"""Copies the reference's one end-to-end example into this folder: its input
files data/db.fasta and data/queries.fasta (byte for byte) and the eight
result lines that README.md:88-97 prints for

    ghostz-gpu db  -i ./data/db.fasta -o exdb
    ghostz-gpu aln -q d -t p -i ./data/queries.fasta -d exdb -o exout

(README.md expands the tab separators, so expected.tsv puts one tab between
the whitespace-separated fields of each line.)

  python tests/golden/readme_example/make_readme_example.py [REFERENCE]
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for name in ("db.fasta", "queries.fasta"):
        shutil.copyfile(os.path.join(ref, "data", name), os.path.join(HERE, name))
    lines = open(os.path.join(ref, "README.md")).read().split("\n")
    block = [ln for ln in lines[87:97] if ln.startswith("query")]
    assert len(block) == 8, block
    with open(os.path.join(HERE, "expected.tsv"), "w") as f:
        for ln in block:
            f.write("\t".join(ln.split()) + "\n")


if __name__ == "__main__":
    main()
