"""ddrs_hotstart_noalloc, the cold start that allocates nothing, against
ddrs_hotstart (its workspace form) and the oracle: two calls in a row, the
second on a smaller network, then the first network again (the scratch keeps
the previous call's words; their epochs keep them from being read), and the
rejection of a plan larger than its scratch. Needs an MI355X."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # collected on the CPU too, skipped there
    pytest.skip("needs a GPU", allow_module_level=True)

from oracle import oracle as O  # noqa: E402
from ddrs_amd import _native as NAT  # noqa: E402
from ddrs_amd import network as NW  # noqa: E402
from ddrs_amd.config import Params  # noqa: E402
from ddrs_amd.sparse import Network, SparseAdjacency  # noqa: E402


def _case(R, seed):
    rows, cols = NW.gen_tree(R, 256, seed)
    net = Network(SparseAdjacency.from_coo(R, rows, cols))
    ws = net.workspace()          # set-up: uploads the plan's schedule
    qp0 = NW.gen_forcing_device(R, 1, seed + 1, True)[0].contiguous()
    return dict(R=R, rows=rows, cols=cols, net=net, ws=ws, qp0=qp0)


def test_noalloc_hotstart_matches_workspace_form():
    print("[case] noalloc hotstart: 300,000 then 20,000 then 300,000 reaches", flush=True)
    cp = Params().to_c()
    stream = torch.cuda.current_stream().cuda_stream
    err = torch.ones(1, dtype=torch.int32, device="cuda:0")
    big, small = _case(300_000, 11), _case(20_000, 13)
    for c in (big, small, big):
        want = torch.empty_like(c["qp0"])
        got = torch.full_like(c["qp0"], float("nan"))
        NAT.check(NAT.lib.ddrs_hotstart(c["ws"].handle, cp, c["qp0"].data_ptr(),
                                        want.data_ptr(), stream), "ddrs_hotstart")
        c["ws"].check()
        NAT.check(NAT.lib.ddrs_hotstart_noalloc(c["net"].handle, cp, c["qp0"].data_ptr(),
                                                got.data_ptr(), err.data_ptr(), stream),
                  "ddrs_hotstart_noalloc")
        torch.cuda.synchronize()
        assert int(err.item()) == 0
        assert torch.equal(got, want)
        ref = O.hotstart(O.make_params(), O.pattern_from_coo(c["R"], c["rows"], c["cols"]),
                         c["qp0"].cpu().numpy())
        assert np.array_equal(got.cpu().numpy(), ref)


def test_noalloc_hotstart_rejects_a_plan_beyond_its_scratch():
    print("[case] noalloc hotstart: capacity", flush=True)
    R = (1 << 22) + 1
    rows, cols = NW.gen_tree(R, 256, 17)
    net = Network(SparseAdjacency.from_coo(R, rows, cols))
    assert net.n > (1 << 22)
    buf = torch.empty(R, device="cuda:0")
    err = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    cp = Params().to_c()
    with pytest.raises(NAT.DdrsError, match="4194304"):
        NAT.check(NAT.lib.ddrs_hotstart_noalloc(net.handle, cp, buf.data_ptr(), buf.data_ptr(),
                                                err.data_ptr(),
                                                torch.cuda.current_stream().cuda_stream),
                  "ddrs_hotstart_noalloc")
