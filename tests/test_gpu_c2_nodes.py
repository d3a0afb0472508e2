"""The join-then-aggregate query as R builds it, through the drop-in nodes:
GroupAgg([attr], s = sum(.vectra_tmp_s), n = n()) <- Project(join columns,
.vectra_tmp_s = value * weight) <- Join(inner | left, fact, dim, fkey = key)
(R/verbs.R:266-284, src/project.c:51-252, src/join.c:262-480).

The GPU GroupAggNode fuses the Project and the GPU JoinNode below it into
one device pass; the oracle runs the same tree node by node.  Keys, group
order and counts bit-exact, sums within 1e-12 * sum |value * weight| of the
group.  Shapes the fused pass does not order (more than 4,096 matches per
probe row, a coerced left join over duplicate keys) run unfused and must
give the same results."""
import numpy as np
import pytest

from helpers import assert_tables_equal, rng, with_na

pytestmark = pytest.mark.gpu


def star(seed, n_fact, n_dim, dup_key_rows=0, fkey_double=False, p_na=0.0):
    r = rng(seed)
    dim_key = r.permutation(n_dim * 3)[:n_dim].astype(np.int64) * 13 - 7
    attr = r.integers(0, 100, n_dim)
    weight = r.random(n_dim)
    ranks = np.minimum(r.zipf(1.1, n_fact) - 1, n_dim - 1)
    fkey = np.where(r.random(n_fact) < 0.1, -(r.integers(1, 10**9, n_fact)), dim_key[ranks])
    if dup_key_rows:
        # the key of Zipf rank 1000 (a few fact rows) on dup_key_rows + 1 dimension rows
        dim_key[1001:1001 + dup_key_rows] = dim_key[1000]
    value = r.lognormal(size=n_fact)
    fact = {"fkey": (fkey.astype(np.float64) if fkey_double else fkey, with_na(r, n_fact, p_na)),
            "value": (value, with_na(r, n_fact, p_na))}
    dim = {"key": dim_key, "attr": (attr, with_na(r, n_dim, p_na)), "weight": (weight, with_na(r, n_dim, p_na))}
    return fact, dim


def tree(nv, engine, fact, dim, kind, temp_dir, rg, pred=None, keys=("attr",), tmp_in=("value", "weight")):
    left = nv.Table(fact, rg_rows=rg).scan()
    if pred is not None:
        left = nv.filter_node(engine, left, pred)
    j = nv.join_node(engine, left, nv.Table(dim).scan(), kind, [(0, 0)])
    if engine == "gpu":
        p = nv.hidden_mutate(j, ".vectra_tmp_s", *tmp_in)
    else:
        p = nv.load_oracle(None).oracle_project_mul_node_create(j, b".vectra_tmp_s", tmp_in[0].encode(),
                                                               tmp_in[1].encode())
        p._keep = j
    aggs = [("s", "sum", ".vectra_tmp_s", 0), ("n", "n", None, 0), ("s_rm", "sum", ".vectra_tmp_s", 1),
            ("m", "mean", ".vectra_tmp_s", 1)]
    return nv.group_agg_node(engine, p, list(keys), aggs, temp_dir)


def abs_scale(nv, fact, dim, kind, pred_mask, got_keys, key_name="attr"):
    """sum |value * weight| per output group, from a host join of the unique
    dimension keys (the tolerance scale)"""
    fk, fv = fact["fkey"][0], fact["fkey"][1]
    val = fact["value"][0]
    dk = dim["key"]
    order = np.argsort(dk)
    pos = np.searchsorted(dk[order], fk)
    pos = np.minimum(pos, len(dk) - 1)
    hit = dk[order][pos] == fk
    if fv is not None:
        hit &= fv
    row = order[pos]
    keep = pred_mask if pred_mask is not None else np.ones(len(fk), bool)
    w = dim["weight"][0][row]
    prod = np.abs(val * w)
    attr_v = dim["attr"][0][row]
    attr_ok = dim["attr"][1][row] if dim["attr"][1] is not None else np.ones(len(fk), bool)
    sel = hit & keep
    scale = {}
    for a in np.unique(attr_v[sel & attr_ok]):
        scale[int(a)] = prod[sel & attr_ok & (attr_v == a)].sum()
    na_scale = prod[sel & ~attr_ok].sum()   # matched rows whose attr is NA (unmatched rows add nothing)
    return np.array([scale.get(int(k), 0.0) if ok else na_scale for k, ok in zip(*got_keys)])


def explain_of(nv, node):
    return "\n".join(nv.explain(node))


@pytest.mark.parametrize("kind", ["inner", "left"])
@pytest.mark.parametrize("temp_dir", ["/tmp", None])
def test_c2_tree_fused(gpu, oracle, kind, temp_dir):
    nv = gpu
    fact, dim = star(7, 400_000, 40_000)
    node = tree(nv, "gpu", fact, dim, kind, temp_dir, 65_536)
    got = nv.collect(node, free=False)
    note = explain_of(nv, node)
    nv.core().vx_node_free(node)
    assert "fused with the Project and Join below" in note, note
    assert "ProjectNode [streaming, hidden mutate]" in note
    exp = nv.collect(tree(nv, "oracle", fact, dim, kind, temp_dir, 65_536))
    scale = abs_scale(nv, fact, dim, kind, None, exp["attr"])
    cnt = np.maximum(exp["n"][0], 1)
    assert_tables_equal(got, exp, {"s": scale, "s_rm": scale, "m": scale / cnt})


@pytest.mark.parametrize("kind", ["inner", "left"])
def test_c2_tree_filtered_probe_with_nas(gpu, oracle, kind):
    """filter(value > t) below the join (a selection vector in every probe
    batch), NA fact keys and values, NA dimension attr and weight"""
    nv = gpu
    fact, dim = star(11, 300_000, 20_000, p_na=0.03)
    pred = ("cmp", ">", ("col", "value"), ("dbl", 0.7))
    got = nv.collect(tree(nv, "gpu", fact, dim, kind, "/tmp", 50_000, pred))
    exp = nv.collect(tree(nv, "oracle", fact, dim, kind, "/tmp", 50_000, pred))
    mask = np.where(fact["value"][1], fact["value"][0] > 0.7, False)
    scale = abs_scale(nv, fact, dim, kind, mask, exp["attr"])
    # groups holding an NA weight or value are NA without na.rm; compare them all
    assert_tables_equal(got, exp, {"s": scale, "s_rm": scale, "m": scale / np.maximum(exp["n"][0], 1)})


@pytest.mark.parametrize("kind", ["inner", "left"])
def test_c2_tree_probe_side_key_and_ungrouped(gpu, oracle, kind):
    """group_by a probe-side column, and an ungrouped summarise, over the
    same fused pass"""
    nv = gpu
    fact, dim = star(13, 200_000, 10_000)
    r = rng(5)
    fact["g"] = r.integers(0, 37, 200_000)
    for keys in (("g",), ()):
        got = nv.collect(tree(nv, "gpu", fact, dim, kind, "/tmp", 30_000, keys=keys))
        exp = nv.collect(tree(nv, "oracle", fact, dim, kind, "/tmp", 30_000, keys=keys))
        for k in keys:
            assert np.array_equal(got[k][0], exp[k][0]) and np.array_equal(got[k][1], exp[k][1])
        assert np.array_equal(got["n"][0], exp["n"][0])
        s_exp, s_got = exp["s"][0], got["s"][0]
        ok = exp["s"][1]
        assert np.array_equal(ok, got["s"][1])
        # every product is positive, so sum |x| of a group is its sum
        assert np.all(np.abs(s_got[ok] - s_exp[ok]) <= 1e-12 * np.abs(s_exp[ok]))


@pytest.mark.parametrize("kind", ["inner", "left"])
def test_c2_tree_unfused_shapes(gpu, oracle, kind):
    """a dimension key with 5,000 rows (more matches per probe row than the
    fused pass orders), and a coerced left join over duplicate keys: the
    node runs the join, the Project and the group-by unfused, same results"""
    nv = gpu
    fact, dim = star(17, 60_000, 8_000, dup_key_rows=5_000)
    node = tree(nv, "gpu", fact, dim, kind, None, 16_384)
    got = nv.collect(node, free=False)
    note = explain_of(nv, node)
    nv.core().vx_node_free(node)
    assert "fused with the Project and Join below" not in note
    exp = nv.collect(tree(nv, "oracle", fact, dim, kind, None, 16_384))
    scale = np.abs(exp["s"][0]) + 1.0
    assert_tables_equal(got, exp, {"s": scale * 1e3, "s_rm": scale * 1e3, "m": scale})


@pytest.mark.parametrize("kind", ["inner", "left"])
def test_c2_tree_coerced_probe_key(gpu, oracle, kind):
    """double fact keys against the int64 dimension key: the build key is
    coerced to double (src/join.c:190-205), the probe key keeps its type;
    and int64 fact keys against a double dimension key: the probe key is
    coerced (inner fused; left over unique keys fused too)"""
    nv = gpu
    fact, dim = star(19, 150_000, 12_000, fkey_double=True)
    got = nv.collect(tree(nv, "gpu", fact, dim, kind, "/tmp", 40_000))
    exp = nv.collect(tree(nv, "oracle", fact, dim, kind, "/tmp", 40_000))
    scale = np.abs(exp["s"][0]) + 1.0
    assert_tables_equal(got, exp, {"s": scale * 1e2, "s_rm": scale * 1e2, "m": scale})
    fact, dim = star(23, 150_000, 12_000)
    dim["key"] = dim["key"].astype(np.float64)
    node = tree(nv, "gpu", fact, dim, kind, "/tmp", 40_000)
    got = nv.collect(node, free=False)
    note = explain_of(nv, node)
    nv.core().vx_node_free(node)
    assert "fused with the Project and Join below" in note
    exp = nv.collect(tree(nv, "oracle", fact, dim, kind, "/tmp", 40_000))
    scale = np.abs(exp["s"][0]) + 1.0
    assert_tables_equal(got, exp, {"s": scale * 1e2, "s_rm": scale * 1e2, "m": scale})
