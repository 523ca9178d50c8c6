"""The host-side training policy of the guided integrator, on the CPU:
updateTrainingStride (integrator.cpp:595-625) and the trainStep batch loop
(integrator.cpp:666-672), checked on hand-worked cases."""
import pytest


def _npmg():
    import npmg

    npmg.lib()
    return npmg


def test_training_stride_cases():
    npmg = _npmg()
    # 2^21 records at a (1, 1) tile for 8 x 65536 per frame: period 4 -> (2, 2)
    assert npmg.training_stride((1, 1), (1, 1), 65536, 8, 1 << 21) == (True, (2, 2))
    # period 5.7: side 2, (2 + 1) * 2 > 5.7 -> (2, 2)
    assert npmg.training_stride((1, 1), (1, 1), 65536, 8, 3000000) == (True, (2, 2))
    # period 8: side 2, then x grows while (x + 1) * 2 <= 8 -> (4, 2)
    assert npmg.training_stride((1, 1), (1, 1), 65536, 8, 1 << 22) == (True, (4, 2))
    # the minimum stride wins over a small period
    assert npmg.training_stride((1, 1), (3, 1), 65536, 8, 1 << 19) == (True, (3, 1))
    # unchanged stride: not accepted
    assert npmg.training_stride((2, 2), (1, 1), 65536, 8, 1 << 19) == (False, (2, 2))
    # a smaller tile is accepted (period 16 / 4 = 4)
    assert npmg.training_stride((4, 4), (1, 1), 65536, 8, 1 << 17) == (True, (2, 2))
    # an equal area in another shape, (2, 3) -> (3, 2): the reference's unsigned
    # sum 1 + (2 - 3) wraps to 0, so it keeps (2, 3)
    assert npmg.training_stride((2, 3), (1, 1), 65536, 8, 1 << 19) == (False, (2, 3))
    # no records: period 0 -> (1, 1); no batch budget: the tile area is bounded
    assert npmg.training_stride((2, 2), (1, 1), 65536, 8, 0) == (True, (1, 1))
    assert npmg.training_stride((1, 1), (1, 1), 0, 8, 1000) == (True, (2048, 2048))


def test_train_batches_cases():
    npmg = _npmg()
    assert npmg.train_batches(600000) == [(0, 524288), (524288, 75648)]  # 75712 rounded down to 128s
    assert npmg.train_batches(100000) == [(0, 99968)]
    assert npmg.train_batches(50000) == []  # below MIN_TRAIN_BATCH_SIZE
    assert npmg.train_batches(10 * 524288) == [(i * 524288, 524288) for i in range(8)]  # batches_per_frame cap
    assert npmg.train_batches(1000, batch_size=256, granularity=1, min_batch=1) == [
        (0, 256), (256, 256), (512, 256), (768, 232)]
    assert npmg.train_batches(1024, batch_size=256, granularity=1, min_batch=1, batches_per_frame=8) == [
        (0, 256), (256, 256), (512, 256), (768, 256)]  # the fifth batch is empty
