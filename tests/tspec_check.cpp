// Host build of the device's template-word resolution (tspec in
// hvm-64_amd/csrc/hvm64_rules.h), checked case by case against the spec
// format that the host compiler writes (hvm64_engine.h): Ref / Num ports and
// P_NONE / P_FREE are literals; a binary tag with slot s in the address field
// is the principal port of the node in slot s, label kept; a Var port with
// 2s+h in the address field is the var cell in half h of slot s; any other
// Red port with r in the address field is register r, as a var endpoint.
// Both register areas (LDS-style and global-style) and both slot strides
// (allocation-ring slices and arena chunks) are covered.  Prints "ok N";
// exits 1 on the first mismatch.
#include <cstdio>
#include <vector>

#include "hvm64_rules.h"

using namespace hvm;

static int checked = 0;

static bool expect(const char *what, u64 got, u64 want) {
  checked++;
  if (got == want) return true;
  std::printf("MISMATCH %s: got %016llx want %016llx\n", what, (unsigned long long)got, (unsigned long long)want);
  return false;
}

int main() {
  const u32 nregs = 12;
  std::vector<u64> area(nregs * 64, 0);
  bool ok = true;
  for (int lds = 0; lds < 2; lds++) {
    for (u64 stride : {(u64)TSLOT_STRIDE, (u64)16}) {
      Interp I;
      I.ipw = 0;
      I.left = 0;
      I.t_lds = lds != 0;
      I.Tl = HVM_TO_LDS(area.data());
      I.Tg = area.data();
      I.tbase = HVM_HEAP_OFF + 0x12340;
      I.tstride = stride;
      // register values (ports, as registers hold): a Var endpoint, a Red
      // (wire) endpoint, a Ctr port, ERA, a Num and an Op port
      const u64 regv[6] = {var_port(HVM_HEAP_OFF + 0x500), wire_port(HVM_HEAP_OFF + 0x708),
                           mkport(T_CTR, 3, HVM_HEAP_OFF + 0x900), P_ERA, numport(0x1234567),
                           mkport(T_OP, 0x2A, HVM_HEAP_OFF + 0xA00)};
      for (u32 r = 0; r < nregs; r++) area[(u64)r * 64] = regv[r % 6];
      // literals
      ok &= expect("ref", tspec(I, mkport(T_REF, 0, def_addr(7))), mkport(T_REF, 0, def_addr(7)));
      ok &= expect("era", tspec(I, P_ERA), P_ERA);
      ok &= expect("num", tspec(I, numport(0xABCDEF3)), numport(0xABCDEF3));
      ok &= expect("none", tspec(I, P_NONE), P_NONE);
      ok &= expect("free", tspec(I, P_FREE), P_FREE);
      // binary nodes of slots 0, 1 and 9, labels kept
      ok &= expect("ctr s0", tspec(I, mkport(T_CTR, 0, 0)), mkport(T_CTR, 0, I.tbase));
      ok &= expect("ctr s9 lab", tspec(I, mkport(T_CTR, 0xFFFE, 9 << 3)), mkport(T_CTR, 0xFFFE, I.tbase + 9 * stride));
      ok &= expect("op s1", tspec(I, mkport(T_OP, 0x21, 1 << 3)), mkport(T_OP, 0x21, I.tbase + stride));
      ok &= expect("swi s4", tspec(I, mkport(T_SWI, 0, 4 << 3)), mkport(T_SWI, 0, I.tbase + 4 * stride));
      // var cells: halves 0 and 1 of slots 0 and 5
      ok &= expect("var s0h0", tspec(I, mkport(T_VAR, 0, 0 << 3)), var_port(I.tbase));
      ok &= expect("var s0h1", tspec(I, mkport(T_VAR, 0, 1 << 3)), var_port(I.tbase + 8));
      ok &= expect("var s5h0", tspec(I, mkport(T_VAR, 0, 10 << 3)), var_port(I.tbase + 5 * stride));
      ok &= expect("var s5h1", tspec(I, mkport(T_VAR, 0, 11 << 3)), var_port(I.tbase + 5 * stride + 8));
      // registers 0 .. nregs-1: endpoints become Var ports, the rest stays
      for (u32 r = 0; r < nregs; r++) {
        const u64 v = regv[r % 6];
        const u64 want = (v & 7) <= T_VAR ? var_port(v & ADDR_MASK) : v;
        char what[32];
        std::snprintf(what, sizeof what, "reg %u", r);
        ok &= expect(what, tspec(I, mkport(T_RED, 0, (u64)r << 3)), want);
      }
    }
  }
  if (!ok) return 1;
  std::printf("ok %d\n", checked);
  return 0;
}
