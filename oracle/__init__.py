"""Oracle: CPU restatement of the reference's cross-sectional analysis path.

TEST INFRASTRUCTURE ONLY. Nothing in the product (the fc package, the
pybind11 modules, libfc_hip.so) imports, links or calls anything in this
directory; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg use it, and only as the checker / the timed CPU baseline.

Modules:
  np_arm      restatement of the reference numpy arm (fc/_cpu_core.py:15-301):
              np_cs_rank, np_parameter_scan, np_rolling_ic, np_stock_corr,
              np_factor_corr and their helpers. Timed as the CPU baseline
              ("port").
  exact       independent definitions used to pin np_arm and the GPU path:
              brute-force ordinal ranks (the contract definition,
              CLAUDE.md:63), exact-rational Spearman IC, per-pair
              np.corrcoef (tests/fixtures/corr_oracle_v1.py:55-88), the
              serial-Kahan Pearson of the HG-2 clause
              (tests/test_adapter_v1.py:96-124), float64 Fraction-exact
              Pearson.
  c/          plain-C stable-sort rank + exact integer IC (rank_oracle.c), a
              fast checker for full-size panels; built by build_c.py into
              oracle/_build/ (git-ignored).

Parity pinning (what checks each function):
  ranks   : frozen anchors (benchmark_corpus/parity_anchors_v1.manifest.json
            case_schema rank_*), brute-force O(N^2) definition, C stable sort.
  IC      : frozen anchors ic_*, exact-rational closed form, np.corrcoef
            Spearman, rolling_ic label fixture.
  corr    : frozen corr corpus (tests/fixtures/corr_corpus_v1.manifest.json),
            anchors corr_*, per-pair np.corrcoef, Fraction-exact Pearson,
            serial-Kahan on the HG-2 bias inputs.
"""
