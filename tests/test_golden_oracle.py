"""The committed fixture (tests/golden/plain_512.json) pins both the synthetic
scene builder (node-array digest) and the oracle's outputs on the plain_512
workload; a change to either shows up here."""
import golden.make_golden as mg


def test_builder_matches_fixture(yx):
    gold = mg.load()
    s = yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512)
    nodes = s.nodes()
    assert nodes.shape[0] == gold["num_nodes"]
    assert mg.node_digest(nodes) == gold["nodes_digest"]


def test_oracle_matches_fixture(yx, orc):
    gold = mg.load()
    s = yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512)
    nodes, root = s.nodes(), s.GetRoot()
    for ent in gold["views"]:
        o = orc.render(nodes, root, ent["pos"], ent["dir"], fov=gold["fov"], w=gold["w"], h=gold["h"],
                       light_mode=ent["light_mode"], light=gold["light"], threads=4)
        assert mg.digest(o) == ent["digest"], f"view {ent['view']} mode {ent['light_mode']}"
        assert o["rays"] == ent["rays"]


def test_oracle_matches_small_fixture(yx, orc):
    """The 64x64 fixture (tests/golden/plain_64.json): every view in all three
    light modes."""
    gold = mg.load(small=True)
    s = yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512)
    assert mg.node_digest(s.nodes()) == gold["nodes_digest"]
    for ent in gold["views"]:
        o = orc.render(s.nodes(), s.GetRoot(), ent["pos"], ent["dir"], fov=gold["fov"], w=gold["w"], h=gold["h"],
                       light_mode=ent["light_mode"], light=gold["light"], threads=4)
        assert mg.digest(o) == ent["digest"], f"view {ent['view']} mode {ent['light_mode']}"


def test_pipeline_oracle_matches_fixture(yx, oracle_mod):
    """The committed pipeline fixture (tests/golden/pipeline_256.json) pins the
    LOD trace, blur and shade restatement on the plain_512 scene."""
    import golden.make_golden_pipeline as mp
    gold = mp.load()
    po = oracle_mod.PipelineOracle()
    s = yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512)
    nodes, root = s.nodes(), s.GetRoot()
    lod = po.lod_colors(nodes, root)
    for ent in gold["views"]:
        lights = oracle_mod.make_lights([dict(pos=ent["pos"], **gold["light"])])
        o = po.pipeline(nodes, root, gold["depth"], ent["pos"], ent["dir"], fov=gold["fov"], w=gold["w"],
                        h=gold["h"], detail_coef=ent["coef"], lights=lights, lod=lod, threads=4)
        assert mp.digest(o) == ent["digest"], f"view {ent['view']} coef {ent['coef']}"
