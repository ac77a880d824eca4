! magudi_hip_input.f90 -- magudi.inp and bc.dat readers of the Fortran host side.
!
! The option dictionary follows magudi's InputHelper (src/InputHelperImpl.f90:142-269):
! one "key = value" per line, '#' starts a comment, values may be quoted; getOption returns
! the default when a key is absent, getRequiredOption stops with a message.  Keys compare
! exactly.  bc.dat rows (src/RegionImpl.f90:175-337): name, type, grid, normal direction,
! iMin iMax jMin jMax kMin kMax, '#' comments.
module magudi_hip_input

  use, intrinsic :: iso_c_binding, only : c_double
  implicit none
  private

  integer, parameter, public :: OPTION_LENGTH = 256

  type :: t_Option
     character(len = OPTION_LENGTH) :: key = "", val = ""
  end type t_Option

  type, public :: t_PatchRow
     character(len = OPTION_LENGTH) :: name = "", type = ""
     integer :: gridIndex = 1, normalDirection = 0, extent(6) = 1
  end type t_PatchRow

  type(t_Option), allocatable :: dict(:)

  interface getOption
     module procedure getOptionString, getOptionReal, getOptionInteger, getOptionLogical
  end interface getOption

  interface getRequiredOption
     module procedure getRequiredString, getRequiredReal, getRequiredInteger
  end interface getRequiredOption

  public :: parseInputFile, getOption, getRequiredOption, hasOption, readBoundaryConditions,   &
       inputError

contains

  subroutine inputError(message)
    ! gracefulExit (src/ErrorHandlerImpl.f90:162-204) of the single-process host
    character(len = *), intent(in) :: message
    write(*, '(2A)') "magudi_hip: ", trim(message)
    stop 1
  end subroutine inputError

  subroutine stripComment(line)
    character(len = *), intent(inout) :: line
    integer :: i
    i = index(line, '#')
    if (i > 0) line(i:) = ' '
  end subroutine stripComment

  function unquote(s) result(v)
    character(len = *), intent(in) :: s
    character(len = OPTION_LENGTH) :: v
    integer :: n
    v = adjustl(s)
    n = len_trim(v)
    if (n >= 2) then
       if ((v(1:1) == '"' .and. v(n:n) == '"') .or. (v(1:1) == "'" .and. v(n:n) == "'")) then
          v = v(2:n - 1)
       end if
    end if
  end function unquote

  subroutine parseInputFile(filename)
    character(len = *), intent(in) :: filename
    character(len = 1024) :: line
    integer :: u, istat, n, i, lineNo
    open(newunit = u, file = trim(filename), action = 'read', status = 'old', iostat = istat)
    if (istat /= 0) call inputError(trim(filename) // ": File not found or permission denied!")
    n = 0
    do
       read(u, '(A)', iostat = istat) line
       if (istat /= 0) exit
       call stripComment(line)
       if (len_trim(line) > 0) n = n + 1
    end do
    if (n == 0) call inputError(trim(filename) // ": File is empty or does not contain any input!")
    if (allocated(dict)) deallocate(dict)
    allocate(dict(n))
    rewind(u)
    i = 0
    lineNo = 0
    do
       read(u, '(A)', iostat = istat) line
       if (istat /= 0) exit
       lineNo = lineNo + 1
       call stripComment(line)
       if (len_trim(line) == 0) cycle
       i = i + 1
       n = index(line, '=')
       if (n == 0) then
          write(line, '(2A,I0,A)') trim(filename), ": Failed to parse input on line ", lineNo, "!"
          call inputError(line)
       end if
       dict(i)%key = adjustl(line(1:n - 1))
       dict(i)%val = unquote(line(n + 1:))
       if (len_trim(dict(i)%key) == 0) then
          write(line, '(2A,I0,A)') trim(filename), ": Empty parameter key on line ", lineNo, "!"
          call inputError(line)
       end if
    end do
    close(u)
  end subroutine parseInputFile

  function lookup(key, val) result(found)
    character(len = *), intent(in) :: key
    character(len = OPTION_LENGTH), intent(out) :: val
    logical :: found
    integer :: i
    found = .false.
    val = ""
    if (.not. allocated(dict)) return
    ! the last definition of a key wins
    do i = size(dict), 1, -1
       if (trim(dict(i)%key) == trim(key)) then
          val = dict(i)%val
          found = .true.
          return
       end if
    end do
  end function lookup

  function hasOption(key) result(found)
    character(len = *), intent(in) :: key
    logical :: found
    character(len = OPTION_LENGTH) :: val
    found = lookup(key, val)
  end function hasOption

  function getOptionString(key, default) result(v)
    character(len = *), intent(in) :: key, default
    character(len = OPTION_LENGTH) :: v
    if (.not. lookup(key, v)) v = default
  end function getOptionString

  function getOptionReal(key, default) result(v)
    character(len = *), intent(in) :: key
    real(c_double), intent(in) :: default
    real(c_double) :: v
    character(len = OPTION_LENGTH) :: s
    integer :: istat
    v = default
    if (.not. lookup(key, s)) return
    read(s, *, iostat = istat) v
    if (istat /= 0) call inputError("option " // trim(key) // ": not a number: " // trim(s))
  end function getOptionReal

  function getOptionInteger(key, default) result(v)
    character(len = *), intent(in) :: key
    integer, intent(in) :: default
    integer :: v
    character(len = OPTION_LENGTH) :: s
    integer :: istat
    v = default
    if (.not. lookup(key, s)) return
    read(s, *, iostat = istat) v
    if (istat /= 0) call inputError("option " // trim(key) // ": not an integer: " // trim(s))
  end function getOptionInteger

  function getOptionLogical(key, default) result(v)
    character(len = *), intent(in) :: key
    logical, intent(in) :: default
    logical :: v
    character(len = OPTION_LENGTH) :: s
    v = default
    if (.not. lookup(key, s)) return
    select case (trim(adjustl(s)))
    case ("true", "TRUE", "True", ".true.", "T", "t", "1", "yes")
       v = .true.
    case ("false", "FALSE", "False", ".false.", "F", "f", "0", "no")
       v = .false.
    case default
       call inputError("option " // trim(key) // ": not a logical value: " // trim(s))
    end select
  end function getOptionLogical

  subroutine getRequiredString(key, v)
    character(len = *), intent(in) :: key
    character(len = *), intent(out) :: v
    character(len = OPTION_LENGTH) :: s
    if (.not. lookup(key, s)) call inputError("Required parameter '" // trim(key) // "' not found!")
    v = s
  end subroutine getRequiredString

  subroutine getRequiredReal(key, v)
    character(len = *), intent(in) :: key
    real(c_double), intent(out) :: v
    if (.not. hasOption(key)) call inputError("Required parameter '" // trim(key) // "' not found!")
    v = getOptionReal(key, 0.0_c_double)
  end subroutine getRequiredReal

  subroutine getRequiredInteger(key, v)
    character(len = *), intent(in) :: key
    integer, intent(out) :: v
    if (.not. hasOption(key)) call inputError("Required parameter '" // trim(key) // "' not found!")
    v = getOptionInteger(key, 0)
  end subroutine getRequiredInteger

  subroutine readBoundaryConditions(filename, rows)
    character(len = *), intent(in) :: filename
    type(t_PatchRow), allocatable, intent(out) :: rows(:)
    character(len = 1024) :: line
    integer :: u, istat, n, i
    open(newunit = u, file = trim(filename), action = 'read', status = 'old', iostat = istat)
    if (istat /= 0) call inputError(trim(filename) // ": File not found or permission denied!")
    n = 0
    do
       read(u, '(A)', iostat = istat) line
       if (istat /= 0) exit
       call stripComment(line)
       if (len_trim(line) > 0) n = n + 1
    end do
    allocate(rows(n))
    rewind(u)
    i = 0
    do
       read(u, '(A)', iostat = istat) line
       if (istat /= 0) exit
       call stripComment(line)
       if (len_trim(line) == 0) cycle
       i = i + 1
       read(line, *, iostat = istat) rows(i)%name, rows(i)%type, rows(i)%gridIndex,           &
            rows(i)%normalDirection, rows(i)%extent
       if (istat /= 0) call inputError(trim(filename) // ": Failed to parse row: " // trim(line))
    end do
    close(u)
  end subroutine readBoundaryConditions

end module magudi_hip_input
