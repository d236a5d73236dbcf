"""Writes ulysses_1994_1998.csv: for every Carrington rotation of 1994-1998 in
the reference's Geliosphere_K0_phi_table.csv, the table row itself, the
Ulysses position on the rotation's start day (heliocentric range, and
colatitude 90 - ecliptic latitude, as Visualization/
prepare_input_based_on_ulysses.py takes them from ulytrj_helio.asc.txt) and
the Ulysses KET P116/P190 proton fluxes (KET_P116_P190_1994-1998.csv)
averaged over the rotation (the 27 days from its start).  Data only; run
from the repository root:

    python3 geliosphere_amd/data/make_ulysses_table.py /root/reference
"""
import csv
import datetime
import os
import re
import sys


def main(ref):
    vis = os.path.join(ref, "Visualization")
    pos = {}
    for line in open(os.path.join(vis, "ulytrj_helio.asc.txt")):
        c = re.split(r"\s+", line.strip())
        if len(c) >= 13:
            pos[(int(c[0]), int(c[1]))] = (float(c[8]), 90.0 - float(c[11]))
    ket = {}
    with open(os.path.join(vis, "KET_P116_P190_1994-1998.csv")) as f:
        for r in csv.DictReader(f):
            ket[(int(r["year"]), int(r["day_of_year"]))] = (float(r["P116_flux"]), float(r["P190_flux"]))
    out = []
    with open(os.path.join(ref, "Geliosphere_K0_phi_table.csv")) as f:
        lines = f.read().splitlines()
    header = lines[0]
    for line in lines[1:]:
        c = line.split(",")
        year, month, day = int(c[5]), int(c[6]), int(c[7])
        if not 1994 <= year <= 1998:
            continue
        d0 = datetime.date(year, month, day)
        doy = d0.timetuple().tm_yday
        if (year, doy) not in pos:
            continue
        vals = []
        for k in range(27):
            d = d0 + datetime.timedelta(days=k)
            v = ket.get((d.year, d.timetuple().tm_yday))
            if v:
                vals.append(v)
        if not vals:
            continue
        r, theta = pos[(year, doy)]
        out.append([line, r, theta, sum(v[0] for v in vals) / len(vals), sum(v[1] for v in vals) / len(vals),
                    len(vals)])
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ulysses_1994_1998.csv")
    with open(dst, "w") as f:
        f.write("# " + header + "\n")
        f.write("table_row;r_au;theta_deg;ket_p116;ket_p190;ket_days\n")
        for row in out:
            f.write("%s;%.4f;%.2f;%.6e;%.6e;%d\n" % tuple(row))
    print(len(out), "rotations ->", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
