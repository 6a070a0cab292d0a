"""Write tests/golden/reference_scalars.json from the reference's own data.

Reads (as JSON/text, nothing executed) the stored PDFs
tests/reference/pdf_zs*.json and the walkthrough notebook's printed outputs
of /root/reference, and keeps the scalar values this repo's oracle is pinned
on, plus the top-level key set of the stored PDF records (the output schema
the command line reproduces); the explorer notebook's printed ingredient
ladder at z_s = 1 (sigma_kappa, clipped variance and skewness, sigma_DL/D_L,
P(ln mu > 1) per model); and the stored PDFs themselves (P_s on the body
grid linspace(-1, 1, 1601), P_s_tail on linspace(1.005, 9, 1600)) as
reference_pdfs.npz.  Run in the build container: python tests/golden/make_golden.py
"""
import json
import os
import re

import numpy as np

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_scalars.json")


def main():
    fx, fxv = {}, {}
    keys = set()
    pdfs = {}
    for name in ("pdf_zs0.5", "pdf_zs1", "pdf_zs3", "pdf_zs10", "pdf_zs0.5_vp", "pdf_zs1_vp"):
        with open(os.path.join(REF, "tests", "reference", name + ".json")) as f:
            d = json.load(f)
        tag = name.replace("pdf_", "").replace(".", "p")
        pdfs[tag + "_P_s"] = np.asarray(d["P_s"], float)
        pdfs[tag + "_P_s_tail"] = np.asarray(d["P_s_tail"], float)
        if name.endswith("_vp"):
            pc = d["pos_closure"]
            fxv[name] = dict(eta=pc["eta"], cells=pc["cells"], groups=pc["groups"],
                             shells=pc["shells"], target_kappa=pc["target_kappa"],
                             target_gamma=pc["target_gamma"], sigma_kappa=d["sigma_kappa"],
                             mean=d["mean"], var=d["var"], skew=d["skew"], sigmaDL=d["sigmaDL"],
                             grid={k: v for k, v in d["grid"].items() if isinstance(v, (int, float))})
            continue
        keys |= set(d)
        fx[name] = dict(Omega_r=d["dark_energy"]["Omega_r"],
                        primordial_growth=d["dark_energy"]["primordial_growth"],
                        sigma_kappa=d["sigma_kappa"], kbar=d["kbar"],
                        N_expl=d["N_expl"], N_fil=d["N_fil"], zs=d["zs"],
                        mean=d["mean"], var=d["var"], skew=d["skew"], sigmaDL=d["sigmaDL"],
                        grid={k: v for k, v in d["grid"].items() if isinstance(v, (int, float))},
                        cosmo=dict(h=d["h"], Om=d["Om"], s8=d["s8"], Ob=d["Ob"], ns=d["ns"],
                                   zeq=d["zeq"]))
    np.savez_compressed(os.path.join(os.path.dirname(OUT), "reference_pdfs.npz"), **pdfs)
    # explorer notebook: the printed ingredient ladder (build_explorer.py:158-180)
    with open(os.path.join(REF, "notebooks", "sgl_pdf_explorer.ipynb")) as f:
        ex = json.load(f)
    ladder = {}
    for c in ex["cells"]:
        for o in c.get("outputs", []):
            t = "".join(o.get("data", {}).get("text/plain", o.get("text", "")))
            if "halos" not in t or "sigma_kappa" not in t:
                continue
            rows = {}
            for line in t.splitlines():
                mm = re.match(r"^(halos|\+ ellipticity|\+ filaments|\+ subhalos|\+ clustering)\s+(.*)$",
                              line)
                if mm:
                    rows.setdefault(mm.group(1), []).extend(float(v) for v in mm.group(2).split()
                                                            if v != "\\")
            for model, v in rows.items():
                ladder[model] = dict(sigma_kappa=v[0], clipped_var=v[1], clipped_skew=v[2],
                                     sigmaDL=v[3], p_tail1=v[4])
    with open(os.path.join(REF, "notebooks", "sgl_analytic_walkthrough.ipynb")) as f:
        nb = json.load(f)
    text = ""
    for c in nb["cells"]:
        for o in c.get("outputs", []):
            if "text" in o:
                text += "".join(o["text"])
    m = re.search(r"sigma_8 \(top hat, the anchor\) = ([0-9.]+)", text)
    m2 = re.search(r"D\(z=1\) = ([0-9.]+),\s+D\(z=3\) = ([0-9.]+)", text)
    m3 = re.search(r"c = ([0-9.]+), r_s = ([0-9.]+) kpc, kappa_s = ([0-9.]+) \(z_s = 1.0\), "
                   r"x_t = r200/r_s = [0-9.]+ \(r_vir/r_s = ([0-9.]+)\)", text)
    nbv = dict(sigma8_tophat=float(m.group(1)), D1=float(m2.group(1)), D3=float(m2.group(2)),
               nfw_1e14_z04=dict(c=float(m3.group(1)), rs_kpc=float(m3.group(2)),
                                 kappa_s=float(m3.group(3)), rvir_over_rs=float(m3.group(4))))
    out = dict(provenance="sgl_analytic v0.2.0 (MIT): tests/reference/pdf_zs*.json and the "
                          "printed outputs of notebooks/sgl_analytic_walkthrough.ipynb cells 3, 7 and "
                          "notebooks/sgl_pdf_explorer.ipynb cell 14 (ingredient ladder)",
               fixtures=fx, fixtures_vp=fxv, walkthrough=nbv, explorer_ladder_zs1=ladder,
               fixture_keys=sorted(keys))
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
