"""The allocation-failure hook on the host (no GPU needed): it takes a
count, rejects a negative one with BGT_ERR_INVALID and a message, and clears
with 0."""
import pytest


def test_hook_arguments():
    import begputhinning_amd as B
    from begputhinning_amd._lib import BGT_ERR_INVALID
    B.debug_fail_device_allocs(3)
    B.debug_fail_device_allocs(0)
    with pytest.raises(B.BGTError) as e:
        B.debug_fail_device_allocs(-1)
    assert e.value.code == BGT_ERR_INVALID and "negative" in str(e.value)
