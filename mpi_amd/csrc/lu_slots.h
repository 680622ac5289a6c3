// lu_slots.h -- jump-table helpers for register tiles indexed by a run-time
// slot: HZ_SLOTS_FROM(start, BODY) runs BODY(b) for every compile-time slot
// b >= start (entered once through a switch, falling through); HZ_SLOT_AT
// runs it for b == start only.  BODY must guard b < CB itself.
#pragma once

#define HZ_SLOTS_FROM(start, BODY) \
    switch (start) { \
    case 0: BODY(0) [[fallthrough]]; \
    case 1: BODY(1) [[fallthrough]]; \
    case 2: BODY(2) [[fallthrough]]; \
    case 3: BODY(3) [[fallthrough]]; \
    case 4: BODY(4) [[fallthrough]]; \
    case 5: BODY(5) [[fallthrough]]; \
    case 6: BODY(6) [[fallthrough]]; \
    case 7: BODY(7) [[fallthrough]]; \
    case 8: BODY(8) [[fallthrough]]; \
    case 9: BODY(9) [[fallthrough]]; \
    case 10: BODY(10) [[fallthrough]]; \
    case 11: BODY(11) [[fallthrough]]; \
    case 12: BODY(12) [[fallthrough]]; \
    case 13: BODY(13) [[fallthrough]]; \
    case 14: BODY(14) [[fallthrough]]; \
    case 15: BODY(15) [[fallthrough]]; \
    case 16: BODY(16) [[fallthrough]]; \
    case 17: BODY(17) [[fallthrough]]; \
    case 18: BODY(18) [[fallthrough]]; \
    case 19: BODY(19) [[fallthrough]]; \
    case 20: BODY(20) [[fallthrough]]; \
    case 21: BODY(21) [[fallthrough]]; \
    case 22: BODY(22) [[fallthrough]]; \
    case 23: BODY(23) [[fallthrough]]; \
    case 24: BODY(24) [[fallthrough]]; \
    case 25: BODY(25) [[fallthrough]]; \
    case 26: BODY(26) [[fallthrough]]; \
    case 27: BODY(27) [[fallthrough]]; \
    case 28: BODY(28) [[fallthrough]]; \
    case 29: BODY(29) [[fallthrough]]; \
    case 30: BODY(30) [[fallthrough]]; \
    case 31: BODY(31) [[fallthrough]]; \
    case 32: BODY(32) [[fallthrough]]; \
    case 33: BODY(33) [[fallthrough]]; \
    case 34: BODY(34) [[fallthrough]]; \
    case 35: BODY(35) [[fallthrough]]; \
    case 36: BODY(36) [[fallthrough]]; \
    case 37: BODY(37) [[fallthrough]]; \
    case 38: BODY(38) [[fallthrough]]; \
    case 39: BODY(39) [[fallthrough]]; \
    case 40: BODY(40) [[fallthrough]]; \
    case 41: BODY(41) [[fallthrough]]; \
    case 42: BODY(42) [[fallthrough]]; \
    case 43: BODY(43) [[fallthrough]]; \
    case 44: BODY(44) [[fallthrough]]; \
    case 45: BODY(45) [[fallthrough]]; \
    case 46: BODY(46) [[fallthrough]]; \
    case 47: BODY(47) [[fallthrough]]; \
    default: break; \
    }

#define HZ_SLOT_AT(start, BODY) \
    switch (start) { \
    case 0: BODY(0) break; \
    case 1: BODY(1) break; \
    case 2: BODY(2) break; \
    case 3: BODY(3) break; \
    case 4: BODY(4) break; \
    case 5: BODY(5) break; \
    case 6: BODY(6) break; \
    case 7: BODY(7) break; \
    case 8: BODY(8) break; \
    case 9: BODY(9) break; \
    case 10: BODY(10) break; \
    case 11: BODY(11) break; \
    case 12: BODY(12) break; \
    case 13: BODY(13) break; \
    case 14: BODY(14) break; \
    case 15: BODY(15) break; \
    case 16: BODY(16) break; \
    case 17: BODY(17) break; \
    case 18: BODY(18) break; \
    case 19: BODY(19) break; \
    case 20: BODY(20) break; \
    case 21: BODY(21) break; \
    case 22: BODY(22) break; \
    case 23: BODY(23) break; \
    case 24: BODY(24) break; \
    case 25: BODY(25) break; \
    case 26: BODY(26) break; \
    case 27: BODY(27) break; \
    case 28: BODY(28) break; \
    case 29: BODY(29) break; \
    case 30: BODY(30) break; \
    case 31: BODY(31) break; \
    case 32: BODY(32) break; \
    case 33: BODY(33) break; \
    case 34: BODY(34) break; \
    case 35: BODY(35) break; \
    case 36: BODY(36) break; \
    case 37: BODY(37) break; \
    case 38: BODY(38) break; \
    case 39: BODY(39) break; \
    case 40: BODY(40) break; \
    case 41: BODY(41) break; \
    case 42: BODY(42) break; \
    case 43: BODY(43) break; \
    case 44: BODY(44) break; \
    case 45: BODY(45) break; \
    case 46: BODY(46) break; \
    case 47: BODY(47) break; \
    default: break; \
    }
